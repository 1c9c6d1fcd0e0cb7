"""Classification of a test FeatureSpace against a training FeatureSpace
(wndcharm/FeatureSpacePrediction.py in the reference):
FeatureSpaceClassification.NewWND5, with the distances, collision mask,
class similarities and marginal probabilities computed on the GPU
(wndgpu_py_wnd5)."""
import numpy as np

from . import _check, _context, _load, _ptr
from .SingleSamplePrediction import SingleSampleClassification


class FeatureSpaceClassification(object):
    """individual_results: one SingleSampleClassification per test sample
    (FeatureSpacePrediction.py:30-120)."""

    def __init__(self, training_set=None, test_set=None, feature_weights=None, name=None, split_number=None):
        self.training_set = training_set
        self.test_set = test_set
        self.feature_weights = feature_weights
        self.name = name
        self.split_number = split_number
        self.individual_results = []

    @classmethod
    def NewWND5(cls, training_set, test_set, feature_weights=None, name=None, split_number=None, quiet=False,
                error_bars=False):
        """WND5 marginal probabilities of every test sample
        (FeatureSpacePrediction.py:641-799): features times weights (ones when
        feature_weights is None), squared euclidean distances to every
        training sample, distances <= DBL_EPSILON masked, similarity d^-5,
        per class the mean similarity; probabilities = class means over their
        sum, predicted label = the most probable class."""
        from .FeatureSpace import FeatureSpace
        from .FeatureWeights import FeatureWeights, FisherFeatureWeights
        if not isinstance(training_set, FeatureSpace):
            raise ValueError('First argument to New must be of type "FeatureSpace", you gave a {0}'.format(
                type(training_set).__name__))
        if not isinstance(test_set, FeatureSpace):
            raise ValueError('Second argument to New must be of type "FeatureSpace", you gave a {0}'.format(
                type(test_set).__name__))
        if feature_weights is not None and not isinstance(feature_weights, FeatureWeights):
            raise ValueError('Third argument to New must be of type "FeatureWeights" or derived class, you gave '
                             'a {0}'.format(type(feature_weights).__name__))
        if test_set.feature_names != training_set.feature_names:
            raise ValueError("Can't classify, features in test set don't match features in training set.")
        if feature_weights is not None and test_set.feature_names != feature_weights.feature_names:
            raise ValueError("Can't classify, features in test set don't match features in weights.")
        n_feats = len(training_set.feature_names)
        if feature_weights is None:
            feature_weights = FisherFeatureWeights(size=n_feats)
            feature_weights.values = np.ones((n_feats,))
            feature_weights.feature_names = training_set.feature_names
        result_set = cls(training_set, test_set, feature_weights, name, split_number)
        tr = np.ascontiguousarray(training_set.data_matrix, np.float64)
        te = np.ascontiguousarray(test_set.data_matrix, np.float64)
        cls_idx = np.ascontiguousarray(training_set._class_indices(), np.int32)
        w = np.ascontiguousarray(np.asarray(feature_weights.values, np.float64))
        C = training_set.num_classes
        nt = te.shape[0]
        prob = np.zeros((nt, C))
        norm = np.zeros(nt)
        pred = np.zeros(nt, np.int64)
        _check(_load().wndgpu_py_wnd5(_context(), _ptr(tr), _ptr(cls_idx), tr.shape[0], tr.shape[1], C, _ptr(w),
                                      _ptr(te), nt, _ptr(prob), _ptr(norm), _ptr(pred)), 'NewWND5')
        names = test_set._contiguous_sample_names or [None] * nt
        truth = test_set._contiguous_ground_truth_labels or [None] * nt
        for q in range(nt):
            r = SingleSampleClassification()
            if pred[q] > 0:
                r.normalization_factor = float(norm[q])
                r.marginal_probabilities = prob[q].copy()
                r.predicted_class = int(pred[q])
                r.predicted_label = training_set.class_names[pred[q] - 1]
            r.name = names[q]
            r.ground_truth_label = truth[q]
            r.split_number = split_number
            result_set.individual_results.append(r)
        return result_set
