"""Feature weights of the wndcharm package (wndcharm/FeatureWeights.py in the
reference): FisherFeatureWeights with NewFromFile, NewFromFeatureSpace (the
Fisher scores computed on the GPU, wndgpu_py_fisher), Threshold and
EliminateZeros, under the reference's names and error behaviour."""
import numpy as np

from . import _check, _context, _load, _ptr


class FeatureWeights(object):
    """A list of feature names and their weights (FeatureWeights.py:38-118)."""

    def __init__(self, name=None, size=None):
        self.name = name
        self.associated_feature_space = None
        if size is not None:
            self.feature_names = [None] * size
            self.values = [None] * size
        else:
            self.feature_names = None
            self.values = None

    def __len__(self):
        return len(self.values) if self.values is not None else 0

    def __getitem__(self, key):
        new = self.__class__(name=self.name)
        new.feature_names = list(self.feature_names[key])
        new.values = list(self.values[key])
        new.associated_feature_space = self.associated_feature_space
        return new


class FisherFeatureWeights(FeatureWeights):
    """Fisher-discriminant weights (FeatureWeights.py:123-306)."""

    @classmethod
    def NewFromFile(cls, weights_filepath):
        """Reads a weights file written by `wndchrm train -vw...` ("value name"
        per line, FeatureWeights.py:140-161)."""
        weights = cls(name=weights_filepath)
        with open(weights_filepath) as fh:
            rows = [line.strip().split(None, 1) for line in fh.read().splitlines() if line.strip()]
        weights.values = [float(v) for v, _ in rows]
        weights.feature_names = [n for _, n in rows]
        return weights

    @classmethod
    def NewFromFeatureSpace(cls, fs):
        """Fisher score per feature of a normalised FeatureSpace
        (FeatureWeights.py:165-236): population mean, per-class means and
        ddof=1 variances, (sum_c (pop - mean_c)^2 / (C-1)) / mean_c(var_c);
        a zero denominator, NaN and +/-inf give 0."""
        if fs.normalized_against is None:
            raise ValueError("Before generating feature weights, call Normalize() of the feature space.")
        data = np.ascontiguousarray(fs.data_matrix, np.float64)
        cls_idx = np.ascontiguousarray(fs._class_indices(), np.int32)
        w = np.zeros(data.shape[1])
        _check(_load().wndgpu_py_fisher(_context(), _ptr(data), _ptr(cls_idx), data.shape[0], data.shape[1],
                                        int(fs.num_classes), _ptr(w)), 'FisherFeatureWeights')
        new_fw = cls()
        new_fw.feature_names = list(fs.feature_names)
        new_fw.values = w.tolist()
        new_fw.associated_feature_space = fs
        return new_fw

    def EliminateZeros(self):
        """A new instance without the features of weight 0 (FeatureWeights.py:239-247)."""
        new_weights = FisherFeatureWeights()
        kept = [(n, v) for n, v in zip(self.feature_names, self.values) if v != 0]
        new_weights.feature_names = [n for n, _ in kept]
        new_weights.values = [v for _, v in kept]
        return new_weights

    def Threshold(self, num_features_to_be_used=None, _all=False):
        """The top n features by weight, in descending order, ties in their
        current order (FeatureWeights.py:250-306).  n: an int in 1..len, a
        float fraction in (0, 1], None for 15%; _all=True reorders every
        feature, _all='non-zero' stops at the first zero weight."""
        if _all:
            num_features_to_be_used = len(self.values)
        elif num_features_to_be_used is None:
            num_features_to_be_used = int(len(self.values) * 0.15)
        elif isinstance(num_features_to_be_used, float):
            if num_features_to_be_used <= 0 or num_features_to_be_used > 1:
                raise ValueError('Choose feature reduction fraction on interval (0,1] (got "{0}"'.format(
                    num_features_to_be_used))
            num_features_to_be_used = int(round(num_features_to_be_used * len(self)))
        elif num_features_to_be_used > len(self.values) or num_features_to_be_used <= 0:
            raise ValueError('Cannot reduce a set of {0} feature weights to requested {1} features.'.format(
                len(self.values), num_features_to_be_used))
        order = sorted(range(len(self.values)), key=lambda i: self.values[i], reverse=True)
        use = order[:num_features_to_be_used]
        if _all is not True:
            for i, k in enumerate(use):
                if self.values[k] == 0:
                    if _all == 'non-zero':
                        use = use[:i]
                        break
                    raise ValueError(
                        "Can't reduce feature weights \"{0}\" to {1} features. Features ranked {2} and below have "
                        "a Fisher score of 0. Request less features. ".format(self.name, num_features_to_be_used, i))
        new_weights = self.__class__()
        new_weights.values = [self.values[k] for k in use]
        new_weights.feature_names = [self.feature_names[k] for k in use]
        new_weights.associated_feature_space = self.associated_feature_space
        return new_weights
