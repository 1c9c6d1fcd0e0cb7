"""One sample's feature vector (wndcharm/FeatureVector.py in the reference):
the members and calls a classification caller uses -- GenerateFeatures
(computed on the GPU by the engine), LoadSigFile, Normalize (on the GPU,
wndgpu_py_normalize) and FeatureReduce -- under the reference's names."""
import numpy as np

from . import StdFeatureComputationPlans, compute_features, read_sig, read_tiff
from .utils import normalize_by_columns


class FeatureVector(object):
    """FeatureVector (FeatureVector.py:99-1100): feature_names, values,
    feature_set_version, and the normalisation bookkeeping."""

    def __init__(self, **kwargs):
        self.name = None
        self.source_filepath = None
        self.original_px_plane = None
        self.ground_truth_label = None
        self.ground_truth_value = None
        self.feature_names = None
        self.values = None
        self.feature_maxima = None
        self.feature_minima = None
        self.feature_means = None
        self.feature_stdevs = None
        self.normalized_against = None
        self.long = None
        self.color = None
        self.feature_set_version = None
        for k, v in kwargs.items():
            if not hasattr(self, k):
                raise AttributeError('No instance variable named "{0}" in class {1}'.format(k, type(self).__name__))
            setattr(self, k, v)
        if self.feature_set_version is None:
            # major 4; minor 1..4 = short, long, short colour, long colour
            minor = (2 if self.long else 1) + (2 if self.color else 0)
            self.feature_set_version = '4.{0}'.format(minor)
        if self.name is None and self.source_filepath is not None:
            self.name = str(self.source_filepath)

    @property
    def num_features(self):
        return len(self.feature_names) if self.feature_names is not None else 0

    def _plan(self):
        if self.long:
            return (StdFeatureComputationPlans.getFeatureSetLongColor() if self.color
                    else StdFeatureComputationPlans.getFeatureSetLong())
        return (StdFeatureComputationPlans.getFeatureSetColor() if self.color
                else StdFeatureComputationPlans.getFeatureSet())

    def GenerateFeatures(self, write_to_disk=False, quiet=True):
        """Computes the plan's features of the source image on the GPU
        (FeatureVector.GenerateFeatures, FeatureVector.py:669-795, whose core
        is FeatureComputationPlanExecutor.run).  The pixels come from
        original_px_plane when set, else from source_filepath."""
        px = self.original_px_plane
        if px is None:
            if self.source_filepath is None:
                raise ValueError('no pixel plane and no source file to compute features from')
            px = read_tiff(self.source_filepath)
        plan = self._plan()
        self.values = compute_features([px], plan)[0]
        self.feature_names = plan.feature_names()
        return self

    def LoadSigFile(self, sigfile_path, quiet=False):
        """Values (and names, unless feature_names is already set, in which
        case the values are taken in that order) from a .sig file
        (FeatureVector.py:971-1100)."""
        t = read_sig(sigfile_path)
        names = list(t.names)
        values = t.values[0].copy()
        major = self.feature_set_version.split('.')[0] if self.feature_set_version else None
        if major is not None and str(t.version) != major:
            raise ValueError('Desired feature set version "{0}" different from "{1}" in file {2}'.format(
                major, t.version, sigfile_path))
        if self.feature_names:
            if self.feature_names != names:
                have = set(names)
                missing = [n for n in self.feature_names if n not in have]
                if missing:
                    raise ValueError('{0} features missing from {1}'.format(len(missing), sigfile_path))
                index = {n: i for i, n in enumerate(names)}
                values = np.array([values[index[n]] for n in self.feature_names])
        else:
            self.feature_names = names
        self.values = values
        return self

    def Normalize(self, reference_features=None, inplace=True, zscore=False, non_real_check=True, quiet=False):
        """Scales the values against a FeatureSpace's minima / maxima (or
        means / stdevs), FeatureVector.py:829-899."""
        if self.normalized_against:
            raise ValueError('{0} "{1}" has already been normalized against {2}.'.format(
                type(self).__name__, self.name, self.normalized_against))
        if not reference_features:
            raise ValueError("Can't normalize {0} \"{1}\" against itself (Normalize() called with blank arg).".format(
                type(self).__name__, self.name))
        if reference_features.feature_names != self.feature_names:
            raise ValueError("Can't normalize {0} \"{1}\" against {2} \"{3}\": Features don't match.".format(
                type(self).__name__, self.name, type(reference_features).__name__, reference_features.name))
        if not reference_features.normalized_against:
            reference_features.Normalize(quiet=quiet, zscore=zscore, non_real_check=non_real_check)
        mins = maxs = means = stdevs = None
        if not zscore:
            mins, maxs = reference_features.feature_minima, reference_features.feature_maxima
        else:
            means, stdevs = reference_features.feature_means, reference_features.feature_stdevs
        fv = np.array(self.values, np.float64) if not inplace or not isinstance(self.values, np.ndarray) \
            else self.values
        r = normalize_by_columns(fv, mins, maxs, means, stdevs, zscore, non_real_check)
        target = self if inplace else self._copy()
        target.values = fv
        target.feature_minima, target.feature_maxima, target.feature_means, target.feature_stdevs = r
        target.normalized_against = reference_features
        return target

    def _copy(self):
        new = FeatureVector(feature_set_version=self.feature_set_version)
        new.__dict__.update(self.__dict__)
        return new

    def FeatureReduce(self, requested_features, inplace=False, quiet=False):
        """The values of requested_features (names, or an object with
        feature_names) in that order (FeatureVector.py:901-970)."""
        try:
            requested_features = requested_features.feature_names
        except AttributeError:
            pass
        index = {n: i for i, n in enumerate(self.feature_names)}
        missing = [n for n in requested_features if n not in index]
        if missing:
            raise ValueError('Feature Reduction error: {0}/{1} requested features are missing'.format(
                len(missing), len(requested_features)))
        order = [index[n] for n in requested_features]
        new = self if inplace else self._copy()
        new.values = np.asarray(self.values, np.float64)[order]
        new.feature_names = list(requested_features)
        for m in ('feature_maxima', 'feature_minima', 'feature_means', 'feature_stdevs'):
            v = getattr(self, m)
            if v is not None:
                setattr(new, m, np.asarray(v)[order])
        if len(order) != len(index):
            new.feature_set_version = '{0}.0'.format(self.feature_set_version.split('.', 1)[0])
        return new
