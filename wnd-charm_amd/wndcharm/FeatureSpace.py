"""A training or test set of feature vectors (wndcharm/FeatureSpace.py in the
reference): the members and calls the classification callers use --
NewFromFitFile, NewFromFeatureVector, Normalize (on the GPU,
wndgpu_py_normalize) and FeatureReduce -- under the reference's names.
Samples are held grouped by class (the reference's data_list views)."""
import numpy as np

from . import read_fit
from .utils import normalize_by_columns


class FeatureSpace(object):
    """FeatureSpace (FeatureSpace.py:55-2119): data_matrix [num_samples,
    num_features], feature_names, class_names, class_sizes, per-class views
    data_list, and the normalisation bookkeeping."""

    def __init__(self, name=None, source_filepath=None, feature_names=None, data_matrix=None, class_names=None,
                 class_sizes=None, sample_names=None, ground_truth_labels=None, feature_set_version=None):
        self.name = name
        self.source_filepath = source_filepath
        self.feature_names = list(feature_names) if feature_names is not None else None
        self.data_matrix = data_matrix
        self.class_names = list(class_names) if class_names is not None else []
        self.class_sizes = list(class_sizes) if class_sizes is not None else []
        self._contiguous_sample_names = list(sample_names) if sample_names is not None else None
        self._contiguous_ground_truth_labels = list(ground_truth_labels) if ground_truth_labels is not None else None
        self.feature_set_version = feature_set_version
        self.feature_maxima = None
        self.feature_minima = None
        self.feature_means = None
        self.feature_stdevs = None
        self.normalized_against = None
        self.num_samples_per_group = 1
        self.interpolation_coefficients = None

    # ---- shape and per-class views
    @property
    def num_samples(self):
        return 0 if self.data_matrix is None else self.data_matrix.shape[0]

    @property
    def num_features(self):
        return 0 if self.feature_names is None else len(self.feature_names)

    @property
    def num_classes(self):
        return len(self.class_names)

    @property
    def data_list(self):
        """per-class views into data_matrix (FeatureSpace.py:107-150)"""
        out, start = [], 0
        for n in self.class_sizes:
            out.append(self.data_matrix[start:start + n])
            start += n
        return out

    def _class_indices(self):
        """class 1..C of every row (rows are grouped by class)"""
        idx = np.zeros(self.num_samples, np.int32)
        start = 0
        for c, n in enumerate(self.class_sizes):
            idx[start:start + n] = c + 1
            start += n
        return idx

    def __str__(self):
        return '<{0} "{1}" n_features={2} n_samples={3} n_classes={4}>'.format(
            type(self).__name__, self.name, self.num_features, self.num_samples, self.num_classes)

    # ---- constructors
    @classmethod
    def NewFromFitFile(cls, pathname, discrete=True, quiet=False):
        """A .fit training set (FeatureSpace.py:788-990): samples of the same
        class made contiguous in class order (samples of no class last),
        file order kept within a class."""
        t = read_fit(pathname)
        order = sorted(range(len(t.classes)), key=lambda i: (t.classes[i] == 0, t.classes[i]))
        cls_sorted = t.classes[order]
        sizes = [int(np.sum(cls_sorted == c)) for c in range(1, t.n_classes + 1)]
        fs = cls(name=str(pathname), source_filepath=pathname, feature_names=t.names,
                 data_matrix=np.ascontiguousarray(t.values[order], np.float64),
                 class_names=t.class_labels[1:t.n_classes + 1], class_sizes=sizes,
                 sample_names=[t.paths[i] for i in order],
                 ground_truth_labels=[t.class_labels[c] if c else 'UNKNOWN' for c in cls_sorted],
                 feature_set_version='{0}.{1}'.format(t.version, t.feature_vec_type))
        return fs

    @classmethod
    def NewFromFeatureVector(cls, fv):
        """A one-sample feature space (FeatureSpace.py:300-330)."""
        fs = cls(name=fv.name, source_filepath=fv.source_filepath, feature_names=fv.feature_names,
                 data_matrix=np.asarray(fv.values, np.float64).reshape(1, -1).copy(),
                 sample_names=[fv.name], ground_truth_labels=[fv.ground_truth_label],
                 feature_set_version=fv.feature_set_version)
        fs.feature_minima, fs.feature_maxima = fv.feature_minima, fv.feature_maxima
        fs.feature_means, fs.feature_stdevs = fv.feature_means, fv.feature_stdevs
        fs.normalized_against = fv.normalized_against
        return fs

    def _copy(self):
        new = type(self)()
        new.__dict__.update(self.__dict__)
        return new

    # ---- transforms
    def Normalize(self, reference_features=None, inplace=True, zscore=False, non_real_check=True, quiet=False):
        """Min-max scaling to 0..100 (or z-scores) against its own ranges or
        a reference feature space's (FeatureSpace.py:627-700)."""
        if self.normalized_against:
            raise ValueError('{0} "{1}" has already been normalized against {2}.'.format(
                type(self).__name__, self.name, self.normalized_against))
        mins = maxs = means = stdevs = None
        if not reference_features:
            against = 'self'
        else:
            if reference_features.feature_names != self.feature_names:
                raise ValueError("Can't normalize {0} \"{1}\" against {2} \"{3}\": Features don't match.".format(
                    type(self).__name__, self.name, type(reference_features).__name__, reference_features.name))
            if not reference_features.normalized_against:
                reference_features.Normalize(quiet=quiet, zscore=zscore, non_real_check=non_real_check)
            against = reference_features
            if not zscore:
                mins, maxs = reference_features.feature_minima, reference_features.feature_maxima
            else:
                means, stdevs = reference_features.feature_means, reference_features.feature_stdevs
        fs = self.data_matrix if inplace else np.copy(self.data_matrix)
        r = normalize_by_columns(fs, mins, maxs, means, stdevs, zscore, non_real_check)
        target = self if inplace else self._copy()
        target.data_matrix = fs
        target.feature_minima, target.feature_maxima, target.feature_means, target.feature_stdevs = r
        target.normalized_against = against
        return target

    def FeatureReduce(self, requested_features, inplace=False, quiet=False):
        """A feature space of the requested features (names, or an object with
        feature_names) in that order (FeatureSpace.py:1564-1640)."""
        try:
            requested_features = requested_features.feature_names
        except AttributeError:
            pass
        index = {n: i for i, n in enumerate(self.feature_names)}
        missing = [n for n in requested_features if n not in index]
        if missing:
            raise ValueError('Feature Reduction error:\n{0} "{1}" is missing {2}/{3} features that were requested '
                             'in the feature reduction list.'.format(type(self).__name__, self.name, len(missing),
                                                                      len(requested_features)))
        order = [index[n] for n in requested_features]
        new = self if inplace else self._copy()
        new.data_matrix = np.ascontiguousarray(self.data_matrix[:, order])
        new.feature_names = list(requested_features)
        new.name = '{0} (feature reduced)'.format(self.name)
        for m in ('feature_maxima', 'feature_minima', 'feature_means', 'feature_stdevs'):
            v = getattr(self, m)
            if v is not None:
                setattr(new, m, np.asarray(v)[order])
        if self.feature_set_version is not None and len(order) != len(index):
            new.feature_set_version = '{0}.0'.format(self.feature_set_version.split('.', 1)[0])
        return new
