"""Classification result of one sample (wndcharm/SingleSamplePrediction.py in
the reference): SingleSampleClassification.NewWND5, computed on the GPU
(wndgpu_py_wnd5)."""


class SingleSampleClassification(object):
    """marginal_probabilities, normalization_factor, predicted_label
    (SingleSamplePrediction.py:215-270); None when a class had no
    non-colliding training sample."""

    def __init__(self):
        self.name = None
        self.marginal_probabilities = None
        self.normalization_factor = None
        self.predicted_label = None
        self.predicted_class = None
        self.predicted_value = None
        self.ground_truth_label = None
        self.ground_truth_value = None
        self.split_number = None
        self.discrete = True

    def __eq__(self, other):
        # the reference's definition: marginal probabilities within 0.005
        for a, b in zip(self.marginal_probabilities, other.marginal_probabilities):
            if abs(a - b) > 0.005:
                return False
        return True

    def __ne__(self, other):
        return not self == other

    @classmethod
    def NewWND5(cls, training_set, feature_weights, test_samp, quiet=False):
        """training_set: FeatureSpace; feature_weights: FeatureWeights;
        test_samp: FeatureVector (SingleSamplePrediction.py:253-262)."""
        from .FeatureSpace import FeatureSpace
        from .FeatureSpacePrediction import FeatureSpaceClassification
        test_set = FeatureSpace.NewFromFeatureVector(test_samp)
        split = FeatureSpaceClassification.NewWND5(training_set, test_set, feature_weights, quiet=quiet)
        return split.individual_results[0]
