"""wndcharm -- Python surface of the MI355X WND-CHARM feature engine.

Mirrors the names of the reference's SWIG module (wndcharm/swig/wndcharm.i:32-40,
Tasks.i:15-35) that callers use on the feature path:

    plan = wndcharm.StdFeatureComputationPlans.getFeatureSetLong()
    vec = wndcharm.DoubleVector(plan.n_features)
    wndcharm.FeatureComputationPlanExecutor(plan).run(px_plane, vec, 0)

(the call sequence of FeatureVector.GenerateFeatures, wndcharm/FeatureVector.py:724-758),
plus batch entry points that take many images per launch.  Everything runs on
the GPU through the C ABI of lib/libwndgpu.so (include/wndgpu.h); there is no
CPU implementation behind these names, and a missing library or GPU raises.
"""
import ctypes
import os

import numpy as np

__all__ = ['ImageMatrix', 'PyImageMatrix', 'DoubleVector', 'FeatureComputationPlan', 'StdFeatureComputationPlans',
           'FeatureComputationPlanExecutor', 'ImageTransform', 'FeatureAlgorithm', 'compute_features',
           'normalize_train', 'normalize_test', 'fisher_weights', 'wnd5', 'WndGpuError', 'library_path',
           'read_tiff', 'write_sig', 'read_sig', 'write_fit', 'read_fit', 'write_weights', 'read_weights',
           'FeatureTable', 'Sampling', 'FP64', 'FP32', 'BF16X3', 'set_graph_mode', 'graph_stats']

_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.environ.get('WNDGPU_LIB', os.path.join(os.path.dirname(_HERE), 'lib', 'libwndgpu.so'))


class WndGpuError(RuntimeError):
    pass


FP64, FP32, BF16X3 = 0, 1, 2  # precisions of the basis contractions (FeatureComputationPlan.set_precision)


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_lib = None


class _Image(ctypes.Structure):
    _fields_ = [('data', _P), ('height', _I), ('width', _I), ('stride', _I), ('dtype', _I)]


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(library_path):
        raise WndGpuError('libwndgpu.so not found at %s: build it with __graft_entry__.build()' % library_path)
    lib = ctypes.CDLL(library_path)
    sig = {
        'wndgpu_create': (_I, [_I, ctypes.POINTER(_P)]),
        'wndgpu_destroy': (None, [_P]),
        'wndgpu_last_error': (ctypes.c_char_p, [_P]),
        'wndgpu_version': (ctypes.c_char_p, []),
        'wndgpu_plan_standard': (_I, [_I, ctypes.POINTER(_P)]),
        'wndgpu_plan_from_groups': (_I, [ctypes.POINTER(ctypes.c_char_p), _I, ctypes.POINTER(_P)]),
        'wndgpu_plan_destroy': (None, [_P]),
        'wndgpu_plan_n_features': (_I, [_P]),
        'wndgpu_plan_feature_name': (ctypes.c_char_p, [_P, _I]),
        'wndgpu_plan_group_offset': (_I, [_P, ctypes.c_char_p]),
        'wndgpu_plan_set_precision': (_I, [_P, _I]),
        'wndgpu_compute': (_I, [_P, _P, ctypes.POINTER(_Image), _I, _P, _P]),
        'wndgpu_compute_device': (_I, [_P, _P, ctypes.POINTER(_Image), _I, _P, _P]),
        'wndgpu_transform': (_I, [_P, ctypes.c_char_p, _P, _I, _I, _P, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
        'wndgpu_transform_image': (_I, [_P, ctypes.c_char_p, ctypes.POINTER(_Image), _P, ctypes.POINTER(_I),
                                        ctypes.POINTER(_I)]),
        'wndgpu_algorithm': (_I, [_P, ctypes.c_char_p, _P, _I, _I, _P, ctypes.POINTER(_I)]),
        'wndgpu_algorithm_n_features': (_I, [ctypes.c_char_p]),
        'wndgpu_normalize_train': (_I, [_P, _P, _I, _I, _P, _P]),
        'wndgpu_normalize_test': (_I, [_P, _P, _I, _P, _P]),
        'wndgpu_fisher': (_I, [_P, _P, _P, _I, _I, _I, _D, _P]),
        'wndgpu_wnd5': (_I, [_P, _P, _P, _I, _I, _I, _P, _P, _I, _P, _P, _P]),
        'wndgpu_wnn': (_I, [_P, _P, _P, _I, _I, _I, _P, _P, _I, _P, _P, _P]),
        'wndgpu_py_normalize': (_I, [_P, _P, _I, _I, _I, _P, _P, _P, _P, _P]),
        'wndgpu_py_fisher': (_I, [_P, _P, _P, _I, _I, _I, _P]),
        'wndgpu_py_wnd5': (_I, [_P, _P, _P, _I, _I, _I, _P, _P, _I, _P, _P, _P]),
        'wndgpu_last_node_times': (_I, [_P, _P, _P, _I]),
        'wndgpu_last_haralick_work': (_I, [_P, _P, _P, _I]),
        'wndgpu_last_cf_open': (_I, [_P, _P, _I]),
        'wndgpu_prefetch_tables': (_I, [_P, _P, _P, _I]),
        'wndgpu_last_call_times': (_I, [_P, _P, _P, _P]),
        'wndgpu_sampling_n_samples': (_I, [_P]),
        'wndgpu_compute_sampled': (_I, [_P, _P, ctypes.POINTER(_Image), _I, _P, _P, _P]),
        'wndgpu_io_error': (ctypes.c_char_p, []),
        'wndgpu_free': (None, [_P]),
        'wndgpu_tiff_read': (_I, [ctypes.c_char_p, ctypes.POINTER(_P), ctypes.POINTER(_Image)]),
        'wndgpu_sig_write': (_I, [ctypes.c_char_p, ctypes.c_char_p, _I, _P, _I, _I, _P, _P, _I]),
        'wndgpu_sig_read': (_I, [ctypes.c_char_p, _I, ctypes.POINTER(_P)]),
        'wndgpu_fit_write': (_I, [ctypes.c_char_p, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P]),
        'wndgpu_fit_read': (_I, [ctypes.c_char_p, ctypes.POINTER(_P)]),
        'wndgpu_table_shape': (_I, [_P, _P, _P, _P, _P, _P]),
        'wndgpu_table_values': (_P, [_P]),
        'wndgpu_table_classes': (_P, [_P]),
        'wndgpu_table_sample_values': (_P, [_P]),
        'wndgpu_table_feature_name': (ctypes.c_char_p, [_P, _I]),
        'wndgpu_table_class_label': (ctypes.c_char_p, [_P, _I]),
        'wndgpu_table_sample_path': (ctypes.c_char_p, [_P, _I]),
        'wndgpu_table_destroy': (None, [_P]),
        'wndgpu_weights_write': (_I, [ctypes.c_char_p, _P, _P, _I]),
        'wndgpu_weights_read': (_I, [ctypes.c_char_p, _P, _I, _P]),
        'wndgpu_set_graph_mode': (_I, [_P, _I]),
        'wndgpu_graph_stats': (_I, [_P, _P, _P, _P, _P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


_ctx = None


def _context():
    """The process-wide context on device WNDGPU_DEVICE (default 0)."""
    global _ctx
    if _ctx is None:
        lib = _load()
        c = _P()
        rc = lib.wndgpu_create(int(os.environ.get('WNDGPU_DEVICE', '0')), ctypes.byref(c))
        if rc != 0:
            raise WndGpuError('wndgpu_create failed (%d): no usable GPU' % rc)
        _ctx = c
    return _ctx


def _check(rc, what):
    if rc != 0:
        msg = _load().wndgpu_last_error(_ctx).decode() if _ctx is not None else ''
        raise WndGpuError('%s failed (%d): %s' % (what, rc, msg))


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


# ---------------------------------------------------------------- pixel plane
class ImageMatrix(object):
    """FP64 row-major pixel plane (ImageMatrix, src/cmatrix.h:164-339)."""

    def __init__(self, pixels=None):
        self._px = None
        self._rgb = None
        if pixels is not None:
            self.set_pixels(pixels)

    def allocate(self, w, h):
        self._px = np.zeros((int(h), int(w)), np.float64)

    def set_pixels(self, a):
        a = np.asarray(a)
        if a.ndim != 2:
            raise ValueError('pixel plane must be 2-D')
        self._px = np.ascontiguousarray(a, dtype=np.float64)
        self._rgb = None

    @property
    def width(self):
        return 0 if self._px is None else self._px.shape[1]

    @property
    def height(self):
        return 0 if self._px is None else self._px.shape[0]

    def data_ptr(self):
        return None if self._px is None else self._px.ctypes.data

    def as_ndarray(self):
        return self._px

    def OpenImage(self, path, downsample=0, bounding_rect=None, mean=0, stdev=0):
        """Loads an 8/16-bit TIFF with 1 or 3 samples per pixel (ImageMatrix::OpenImage,
        cmatrix.cpp:252-294, through LoadTIFF 85-203); returns 1 on success, 0 on
        failure like the reference.  A colour image keeps its RGB samples (the
        colour plans read them); its pixel plane is RGB2GRAY, computed on the GPU."""
        try:
            a = read_tiff(path)
        except WndGpuError:
            return 0
        if bounding_rect is not None:
            x, y, w, h = bounding_rect
            a = a[y:y + h, x:x + w]
        if a.ndim == 3:
            self.set_rgb(a)
        else:
            self.set_pixels(a.astype(np.float64))
        return 1

    def set_rgb(self, a):
        """an (h, w, 3) uint8 / uint16 colour image: the gray plane is its RGB2GRAY"""
        a = np.ascontiguousarray(a)
        if a.ndim != 3 or a.shape[2] != 3 or a.dtype not in (np.uint8, np.uint16):
            raise ValueError('colour images are (h, w, 3) uint8 or uint16 arrays')
        self._px = ImageTransform('Color Transform').execute(a)
        self._rgb = a

    def source(self):
        """what the engine is given: the RGB samples of a colour image, else the plane"""
        return self._rgb if self._rgb is not None else self._px


class PyImageMatrix(ImageMatrix):
    """numpy view helper of the reference (wndcharm/PyImageMatrix.py:41-63)."""


class DoubleVector(object):
    """std::vector<double> stand-in of the SWIG module (Tasks.i:15-17)."""

    def __init__(self, n=0):
        self._a = np.zeros(int(n), np.float64)

    def __len__(self):
        return len(self._a)

    def __getitem__(self, i):
        return self._a[i]

    def __setitem__(self, i, v):
        self._a[i] = v

    def __iter__(self):
        return iter(self._a.tolist())

    def resize(self, n):
        a = np.zeros(int(n), np.float64)
        k = min(n, len(self._a))
        a[:k] = self._a[:k]
        self._a = a

    def as_ndarray(self):
        return self._a


# ---------------------------------------------------------------- plans
class FeatureComputationPlan(object):
    """FeatureComputationPlan (src/Tasks.h:272-300, src/Tasks.cpp:118-168)."""

    def __init__(self, name='', _handle=None, _vec_type=0):
        self.name = name
        self._groups = []
        self._handle = _handle
        self.feature_vec_type = _vec_type
        self._final = _handle is not None

    def add(self, group_name):
        if self._final:
            raise WndGpuError('Attempt to call FeatureComputationPlan::add() to a finalized plan')
        if group_name in self._groups:
            raise WndGpuError('Calling FeatureComputationPlan::add() with duplicate FeatureGroup')
        self._groups.append(group_name)

    def finalize(self):
        if self._final:
            return
        lib = _load()
        arr = (ctypes.c_char_p * len(self._groups))(*[g.encode() for g in self._groups])
        h = _P()
        rc = lib.wndgpu_plan_from_groups(arr, len(self._groups), ctypes.byref(h))
        if rc != 0:
            raise WndGpuError('unknown or duplicate feature group in %r' % (self._groups,))
        self._handle = h
        self._final = True

    def isFinalized(self):
        return self._final

    @property
    def handle(self):
        self.finalize()
        return self._handle

    @property
    def n_features(self):
        return _load().wndgpu_plan_n_features(self.handle)

    def getFeatureNameByIndex(self, i):
        return _load().wndgpu_plan_feature_name(self.handle, int(i)).decode()

    def feature_names(self):
        return [self.getFeatureNameByIndex(i) for i in range(self.n_features)]

    def getFGoffset(self, group_name):
        return _load().wndgpu_plan_group_offset(self.handle, group_name.encode())

    def set_precision(self, precision):
        """FP64 (0, the plain form), FP32 (1) or BF16X3 (2) for the basis contractions
        of Chebyshev, Zernike, Chebyshev-Fourier Coefficients and Gabor Textures
        (BF16X3 on the matrix cores for the GEMM-shaped Chebyshev and Gabor,
        FP32 for the other two); the standard plans are shared, so set it on a
        plan of your own"""
        if self._handle is not None and any(self is p for p in StdFeatureComputationPlans._cache.values()):
            raise WndGpuError('set the precision on a plan of your own, not a shared standard plan')
        _check(_load().wndgpu_plan_set_precision(self.handle, int(precision)), 'set_precision')
        self.precision = int(precision)


class StdFeatureComputationPlans(object):
    """StdFeatureComputationPlans (src/Tasks.h:304-377): process singletons."""
    _cache = {}

    @classmethod
    def _get(cls, s, vec_type, name):
        if s not in cls._cache:
            h = _P()
            _check(_load().wndgpu_plan_standard(s, ctypes.byref(h)), 'wndgpu_plan_standard')
            cls._cache[s] = FeatureComputationPlan(name, _handle=h, _vec_type=vec_type)
        return cls._cache[s]

    @classmethod
    def getFeatureSet(cls):
        return cls._get(0, 1, 'Standard Feature Set')

    @classmethod
    def getFeatureSetLong(cls):
        return cls._get(1, 2, 'Long Feature Set')

    @classmethod
    def getFeatureSetColor(cls):
        return cls._get(2, 3, 'Color Feature Set')

    @classmethod
    def getFeatureSetLongColor(cls):
        return cls._get(3, 4, 'Long Color Feature Set')


def _as_image_struct(a):
    if a.ndim == 3:  # (h, w, 3) interleaved RGB, 8 or 16 bits per sample
        if a.shape[2] != 3 or a.dtype not in (np.uint8, np.uint16):
            raise ValueError('colour images are (h, w, 3) uint8 or uint16 arrays')
        a = np.ascontiguousarray(a)
        dt = 4 if a.dtype == np.uint8 else 5
        return a, _Image(ctypes.c_void_p(a.ctypes.data), a.shape[0], a.shape[1], a.shape[1], dt)
    if a.ndim != 2:
        raise ValueError('images are 2-D arrays or (h, w, 3) RGB arrays')
    if a.dtype == np.uint16:
        dt = 0
    elif a.dtype == np.float32:
        dt = 1
    elif a.dtype == np.float64:
        dt = 2
    elif a.dtype == np.uint8:
        dt = 3
    else:
        a = a.astype(np.float64)
        dt = 2
    a = np.ascontiguousarray(a)
    return a, _Image(ctypes.c_void_p(a.ctypes.data), a.shape[0], a.shape[1], a.shape[1], dt)


class Sampling(ctypes.Structure):
    """How each image is sampled before the features (wndchrm's pre-processing
    and sampling options): crop to roi = (x, y, w, h), downsample in percent,
    normalise to mean (and stdev), then rotations x tiles = (tiles_x, tiles_y)
    samples, rotation outer, tile y inner (TrainingSet.cpp:1425-1480)."""
    _fields_ = [('roi_x', _I), ('roi_y', _I), ('roi_w', _I), ('roi_h', _I), ('downsample', _I), ('mean', _D),
                ('stdev', _D), ('rotations', _I), ('tiles_x', _I), ('tiles_y', _I)]

    def __init__(self, roi=None, downsample=0, mean=0.0, stdev=0.0, rotations=1, tiles=(1, 1)):
        x, y, w, h = roi if roi is not None else (-1, -1, 0, 0)
        super(Sampling, self).__init__(x, y, w, h, int(downsample), float(mean), float(stdev), int(rotations),
                                       int(tiles[0]), int(tiles[1]))

    @property
    def n_samples(self):
        n = _load().wndgpu_sampling_n_samples(ctypes.byref(self))
        if n < 0:
            raise WndGpuError('rotations must be 1..4 and tiles at least 1')
        return n


def compute_features(images, plan=None, return_status=False, sampling=None):
    """Features of a batch of 2-D arrays (uint8/uint16/float32/float64) or
    (h, w, 3) RGB arrays (uint8/uint16; the colour plans need these), sizes
    may differ, in one pass of the DAG on the GPU.  Returns [n][n_features]
    float64; rows of rejected images are NaN (status != 0).  With a Sampling,
    each image gives n_samples rows (sample k of image i at i * n_samples + k)."""
    if plan is None:
        plan = StdFeatureComputationPlans.getFeatureSetLong()
    keep = []
    structs = (_Image * len(images))()
    for i, im in enumerate(images):
        a, s = _as_image_struct(np.asarray(im))
        keep.append(a)
        structs[i] = s
    nf = plan.n_features
    status = np.zeros(len(images), np.int32)
    if sampling is not None:
        # rows: sample k of image i at i * n_samples + k
        out = np.full((len(images) * sampling.n_samples, nf), np.nan)
        rc = _load().wndgpu_compute_sampled(_context(), plan.handle, structs, len(images), ctypes.byref(sampling),
                                            _ptr(out), _ptr(status))
    else:
        out = np.full((len(images), nf), np.nan)
        rc = _load().wndgpu_compute(_context(), plan.handle, structs, len(images), _ptr(out), _ptr(status))
    if rc not in (0, -4):
        _check(rc, 'wndgpu_compute')
    if return_status:
        return out, status
    if rc == -4:
        raise WndGpuError('images rejected, status=%s' % status.tolist())
    return out


class FeatureComputationPlanExecutor(object):
    """FeatureComputationPlanExecutor (src/Tasks.h:335-357, Tasks.cpp:275-295)."""

    def __init__(self, plan):
        self.plan = plan

    def run(self, source_mat, feature_mat, dest_row):
        px = source_mat.source() if hasattr(source_mat, 'source') else np.asarray(source_mat)
        if px is None:
            raise WndGpuError('Attempt to execute a FeatureComputationPlan with an empty ImageMatrix')
        nf = self.plan.n_features
        row = compute_features([px], self.plan)[0]
        dst = feature_mat.as_ndarray() if hasattr(feature_mat, 'as_ndarray') else feature_mat
        if len(dst) < nf * (dest_row + 1):
            raise WndGpuError('feature matrix too small for row %d' % dest_row)
        dst[nf * dest_row:nf * (dest_row + 1)] = row


class ImageTransform(object):
    """ImageTransform::execute by registered name (src/ImageTransforms.h:45-51)."""

    def __init__(self, name):
        self.name = name

    def execute(self, matrix_in, matrix_out=None):
        """matrix_in: a gray plane (ImageMatrix or 2-D array), or an (h, w, 3)
        uint8 / uint16 RGB array, which "Color" and "Hue" need."""
        a = matrix_in.as_ndarray() if hasattr(matrix_in, 'as_ndarray') else np.asarray(matrix_in)
        oh, ow = _I(), _I()
        if a.ndim == 3:
            a, im = _as_image_struct(a)
            h, w = a.shape[:2]
            out = np.zeros((2 * ((h + 9) // 2)) * (2 * ((w + 9) // 2)))
            _check(_load().wndgpu_transform_image(_context(), self.name.encode(), ctypes.byref(im), _ptr(out),
                                                  ctypes.byref(oh), ctypes.byref(ow)), 'ImageTransform %s' % self.name)
        else:
            a = np.ascontiguousarray(a, dtype=np.float64)
            h, w = a.shape
            out = np.zeros((2 * ((h + 9) // 2)) * (2 * ((w + 9) // 2)))
            _check(_load().wndgpu_transform(_context(), self.name.encode(), _ptr(a), h, w, _ptr(out),
                                            ctypes.byref(oh), ctypes.byref(ow)), 'ImageTransform %s' % self.name)
        r = out[:oh.value * ow.value].reshape(oh.value, ow.value)
        if matrix_out is not None:
            matrix_out.set_pixels(r)
            return matrix_out
        return r


class FeatureAlgorithm(object):
    """FeatureAlgorithm::execute by registered name (src/FeatureAlgorithms.h:280-290)."""

    def __init__(self, name):
        self.name = name
        self.n_features = _load().wndgpu_algorithm_n_features(name.encode())
        if self.n_features < 0:
            raise WndGpuError('unknown feature algorithm %r' % name)

    def execute(self, matrix_in):
        a = np.ascontiguousarray(matrix_in.as_ndarray() if hasattr(matrix_in, 'as_ndarray') else matrix_in,
                                 dtype=np.float64)
        out = np.zeros(self.n_features)
        n = _I()
        _check(_load().wndgpu_algorithm(_context(), self.name.encode(), _ptr(a), a.shape[0], a.shape[1], _ptr(out),
                                        ctypes.byref(n)), 'FeatureAlgorithm %s' % self.name)
        return out


# ---------------------------------------------------------------- classifier
def normalize_train(data):
    """TrainingSet::normalize: returns (normalised copy, mins, maxs)."""
    d = np.ascontiguousarray(np.array(data, np.float64))
    n, f = d.shape
    mins, maxs = np.zeros(f), np.zeros(f)
    _check(_load().wndgpu_normalize_train(_context(), _ptr(d), n, f, _ptr(mins), _ptr(maxs)), 'normalize_train')
    return d, mins, maxs


def normalize_test(vec, mins, maxs):
    v = np.ascontiguousarray(np.array(vec, np.float64))
    mins = np.ascontiguousarray(mins, np.float64)
    maxs = np.ascontiguousarray(maxs, np.float64)
    _check(_load().wndgpu_normalize_test(_context(), _ptr(v), len(v), _ptr(mins), _ptr(maxs)), 'normalize_test')
    return v


def fisher_weights(data, classes, n_classes, used_fraction=0.15):
    d = np.ascontiguousarray(data, np.float64)
    c = np.ascontiguousarray(classes, np.int32)
    if d.ndim != 2 or c.shape != (d.shape[0],):
        raise WndGpuError('fisher_weights: %d samples and %d class labels' % (d.shape[0], c.size))
    w = np.zeros(d.shape[1])
    _check(_load().wndgpu_fisher(_context(), _ptr(d), _ptr(c), d.shape[0], d.shape[1], int(n_classes),
                                 float(used_fraction), _ptr(w)), 'fisher')
    return w


def wnd5(train, classes, n_classes, weights, test):
    """classify2 for a batch of normalised test vectors: (probabilities,
    norm_factor, predicted class 1..n_classes)."""
    tr = np.ascontiguousarray(train, np.float64)
    c = np.ascontiguousarray(classes, np.int32)
    w = np.ascontiguousarray(weights, np.float64)
    te = np.ascontiguousarray(np.atleast_2d(test), np.float64)
    if tr.ndim != 2 or c.shape != (tr.shape[0],) or w.shape != (tr.shape[1],) or te.shape[1] != tr.shape[1]:
        raise WndGpuError('wnd5: shapes of train %s, classes %s, weights %s and test %s disagree'
                          % (tr.shape, c.shape, w.shape, te.shape))
    nt = te.shape[0]
    prob = np.zeros((nt, n_classes))
    norm = np.zeros(nt)
    pred = np.zeros(nt, np.int64)
    _check(_load().wndgpu_wnd5(_context(), _ptr(tr), _ptr(c), tr.shape[0], tr.shape[1], int(n_classes), _ptr(w),
                               _ptr(te), nt, _ptr(prob), _ptr(norm), _ptr(pred)), 'wnd5')
    return prob, norm, pred


def wnn(train, classes, n_classes, weights, test):
    """WNNclassify (weighted nearest neighbour) for a batch of normalised
    test vectors: (probabilities, norm_factor, predicted class 1..n_classes)."""
    tr = np.ascontiguousarray(train, np.float64)
    c = np.ascontiguousarray(classes, np.int32)
    w = np.ascontiguousarray(weights, np.float64)
    te = np.ascontiguousarray(np.atleast_2d(test), np.float64)
    if tr.ndim != 2 or c.shape != (tr.shape[0],) or w.shape != (tr.shape[1],) or te.shape[1] != tr.shape[1]:
        raise WndGpuError('wnn: shapes of train %s, classes %s, weights %s and test %s disagree'
                          % (tr.shape, c.shape, w.shape, te.shape))
    nt = te.shape[0]
    prob = np.zeros((nt, n_classes))
    norm = np.zeros(nt)
    pred = np.zeros(nt, np.int64)
    _check(_load().wndgpu_wnn(_context(), _ptr(tr), _ptr(c), tr.shape[0], tr.shape[1], int(n_classes), _ptr(w),
                              _ptr(te), nt, _ptr(prob), _ptr(norm), _ptr(pred)), 'wnn')
    return prob, norm, pred


def set_graph_mode(mode):
    """0: every batch issues its launches; 1: repeated batch signatures are
    captured once into a hipGraph and replayed; 2: also pick the faster of
    the two launch schedules by timing both (wndgpu_set_graph_mode)."""
    _check(_load().wndgpu_set_graph_mode(_context(), int(mode)), 'set_graph_mode')


def graph_stats():
    """(captures, replays, graphs cached, schedule of the last graph or -1)"""
    cap, rep = ctypes.c_long(), ctypes.c_long()
    cached, sched = _I(), _I()
    _check(_load().wndgpu_graph_stats(_context(), ctypes.byref(cap), ctypes.byref(rep), ctypes.byref(cached),
                                      ctypes.byref(sched)), 'graph_stats')
    return cap.value, rep.value, cached.value, sched.value


def last_node_times():
    """(name, ms) per launch of the last compute, when WNDGPU_TIMING is set."""
    lib = _load()
    ms = np.zeros(1024)
    names = (ctypes.c_char_p * 1024)()
    k = lib.wndgpu_last_node_times(_context(), _ptr(ms), names, 1024)
    return [(names[i].decode(), float(ms[i])) for i in range(k)]


def last_haralick_work():
    """(tones [k], schur_flops [k, 4]) of the Haralick entries of the last compute."""
    lib = _load()
    n = lib.wndgpu_last_haralick_work(_context(), None, None, 0)
    tones = np.zeros(max(n, 1), np.int32)
    flops = np.zeros((max(n, 1), 4))
    lib.wndgpu_last_haralick_work(_context(), _ptr(tones), _ptr(flops), n)
    return tones[:n], flops[:n]


def prefetch_tables(images, plan=None):
    """Queues the size-dependent tables a later compute_features(images,
    plan) will need on the library's host threads and returns at once (the
    pixels are not read)."""
    if plan is None:
        plan = StdFeatureComputationPlans.getFeatureSetLong()
    keep = []
    structs = (_Image * len(images))()
    for i, im in enumerate(images):
        a, st = _as_image_struct(np.asarray(im))
        keep.append(a)
        structs[i] = st
    _check(_load().wndgpu_prefetch_tables(_context(), plan.handle, structs, len(images)), 'prefetch_tables')


def last_call_times():
    """(host_ms, table_ms, device_ms) of the last compute call."""
    v = [ctypes.c_double() for _ in range(3)]
    _check(_load().wndgpu_last_call_times(_context(), *[ctypes.byref(x) for x in v]), 'last_call_times')
    return tuple(x.value for x in v)


def last_cf_open():
    """[k] int32 of the Chebyshev-Fourier entries of the last compute: 1 where
    the matrix-core histogram was left uncertified and the entry was
    recomputed in the reference's order."""
    lib = _load()
    n = lib.wndgpu_last_cf_open(_context(), None, 0)
    out = np.zeros(max(n, 1), np.int32)
    lib.wndgpu_last_cf_open(_context(), _ptr(out), n)
    return out[:n]


# ---------------------------------------------------------------- files
def _io_check(rc, what):
    if rc != 0:
        raise WndGpuError('%s: %s' % (what, _load().wndgpu_io_error().decode(errors='replace')))


def _cstrs(items):
    arr = (ctypes.c_char_p * max(len(items), 1))(*[str(x).encode() for x in items])
    return arr


def read_tiff(path):
    """A TIFF image as ImageMatrix::LoadTIFF reads it (cmatrix.cpp:85-203):
    (h, w) uint8 / uint16, or (h, w, 3) for 3 samples per pixel."""
    lib = _load()
    buf = _P()
    im = _Image()
    _io_check(lib.wndgpu_tiff_read(os.fsencode(path), ctypes.byref(buf), ctypes.byref(im)), 'read_tiff')
    try:
        dt = {0: np.uint16, 3: np.uint8, 4: np.uint8, 5: np.uint16}[im.dtype]
        ch = 3 if im.dtype in (4, 5) else 1
        n = im.height * im.width * ch
        raw = ctypes.string_at(buf, n * np.dtype(dt).itemsize)
        a = np.frombuffer(raw, dt).copy()
    finally:
        lib.wndgpu_free(buf)
    return a.reshape((im.height, im.width, 3) if ch == 3 else (im.height, im.width))


def write_sig(path, values, names=None, image_path='', sample_class=0, sample_value=None, version=4, vec_type=0):
    """signatures::SaveToFile (signatures.cpp:227-261): byte-identical .sig text"""
    v = np.ascontiguousarray(values, dtype=np.float64)
    sv = ctypes.c_double(sample_value) if sample_value is not None else None
    nm = _cstrs(names) if names is not None else None
    _io_check(_load().wndgpu_sig_write(os.fsencode(path), image_path.encode(), int(sample_class),
                                       ctypes.byref(sv) if sv is not None else None, int(version), int(vec_type),
                                       _ptr(v), nm, len(v)), 'write_sig')


class FeatureTable(object):
    """a .sig (one sample) or .fit (training set) as read from its file"""

    def __init__(self, handle):
        lib = _load()
        shape = [_I() for _ in range(5)]
        lib.wndgpu_table_shape(handle, *[ctypes.byref(x) for x in shape])
        ns, nf, nc, self.version, self.feature_vec_type = [x.value for x in shape]
        self.n_classes = nc
        vals = lib.wndgpu_table_values(handle)
        self.values = (np.ctypeslib.as_array(ctypes.cast(vals, ctypes.POINTER(_D)), (ns * nf,)).copy()
                       if ns * nf else np.zeros(0)).reshape(ns, nf)
        cl = lib.wndgpu_table_classes(handle)
        self.classes = np.ctypeslib.as_array(ctypes.cast(cl, ctypes.POINTER(_I)), (ns,)).copy() if ns else \
            np.zeros(0, np.int32)
        sv = lib.wndgpu_table_sample_values(handle)
        self.sample_values = np.ctypeslib.as_array(ctypes.cast(sv, ctypes.POINTER(_D)), (ns,)).copy() if ns else \
            np.zeros(0)
        self.names = [lib.wndgpu_table_feature_name(handle, i).decode() for i in range(nf)]
        labels = []
        i = 0
        while True:
            lab = lib.wndgpu_table_class_label(handle, i)
            if lab is None:
                break
            labels.append(lab.decode())
            i += 1
        self.class_labels = labels
        self.paths = [lib.wndgpu_table_sample_path(handle, i).decode() for i in range(ns)]


def read_sig(path, continuous=False):
    """signatures::LoadFromFilep (signatures.cpp:277-304): a FeatureTable with one sample"""
    lib = _load()
    h = _P()
    _io_check(lib.wndgpu_sig_read(os.fsencode(path), int(bool(continuous)), ctypes.byref(h)), 'read_sig')
    try:
        return FeatureTable(h)
    finally:
        lib.wndgpu_table_destroy(h)


def read_fit(path):
    """TrainingSet::ReadFromFile (TrainingSet.cpp:402-480)"""
    lib = _load()
    h = _P()
    _io_check(lib.wndgpu_fit_read(os.fsencode(path), ctypes.byref(h)), 'read_fit')
    try:
        return FeatureTable(h)
    finally:
        lib.wndgpu_table_destroy(h)


def write_fit(path, values, classes, feature_names, class_labels, sample_paths=None, sample_values=None,
              version=4, vec_type=0):
    """TrainingSet::SaveToFile (TrainingSet.cpp:339-380); class_labels has n_classes + 1
    entries (entry 0 is the unknown class's label)"""
    v = np.ascontiguousarray(values, dtype=np.float64)
    ns, nf = v.shape
    c = np.ascontiguousarray(classes, dtype=np.int32)
    sv = np.ascontiguousarray(sample_values, dtype=np.float64) if sample_values is not None else None
    _io_check(_load().wndgpu_fit_write(os.fsencode(path), len(class_labels) - 1, int(version), int(vec_type), ns, nf,
                                       _ptr(v), _ptr(c), _ptr(sv) if sv is not None else None,
                                       _cstrs(feature_names), _cstrs(class_labels),
                                       _cstrs(sample_paths) if sample_paths is not None else None), 'write_fit')


def write_weights(path, weights, names):
    """TrainingSet::SaveWeightVector (TrainingSet.cpp:594-606)"""
    w = np.ascontiguousarray(weights, dtype=np.float64)
    _io_check(_load().wndgpu_weights_write(os.fsencode(path), _ptr(w), _cstrs(names), len(w)), 'write_weights')


def read_weights(path, cap=100000):
    """TrainingSet::LoadWeightVector (TrainingSet.cpp:613-642) values (negative read as 0)"""
    w = np.zeros(cap)
    n = _I()
    _io_check(_load().wndgpu_weights_read(os.fsencode(path), _ptr(w), cap, ctypes.byref(n)), 'read_weights')
    return w[:n.value].copy()
