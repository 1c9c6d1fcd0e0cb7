"""Differentiable form of the feature path: torch.autograd over the adjoint kernels.

The reference computes features and WND5 probabilities but no derivatives.
This module gives the gradient of the WND5 marginal probabilities with
respect to the pixels of FP64 device planes, through hand-written adjoint
kernels behind the C ABI (``wndgpu_ad_*`` in include/wndgpu.h):

* transforms  "Fourier" (ImageMatrix::fft2, src/cmatrix.cpp:1107-1200) and
  "Chebyshev" (ImageMatrix::ChebyshevTransform, src/cmatrix.cpp:1203-1222);
* features    "Zernike Coefficients" (mb_zernike2D,
  src/textures/zernike/zernike.cpp:207-341, 72 values, identical to the
  feature path's), "Chebyshev Coefficients (raw)", the 20 x 20
  coefficients that "Chebyshev Coefficients" histograms
  (src/FeatureAlgorithms.cpp:73-93), and "Chebyshev-Fourier Coefficients
  (raw)", the 23 x 23 magnitudes that "Chebyshev-Fourier Coefficients"
  histograms, after its Downsample above 300 x 300 (src/FeatureAlgorithms.cpp:54-69),
  "Comb Moments (raw)", the mean / std / skew / kurtosis of the 4 x 20
  pixel lines that "Comb Moments" histograms (src/FeatureAlgorithms.cpp:184-200),
  and "Gabor Textures (raw)", the means of the 8 energy maps that "Gabor
  Textures" thresholds (src/FeatureAlgorithms.cpp:422-434; the maps
  themselves from gabor_energy());
* WND5        normalisation against the training set
  (signatures::normalize, src/wndchrm_src/signatures.cpp:196-215) followed by
  TrainingSet::classify2 (src/wndchrm_src/TrainingSet.cpp:2429-2586).

Feature groups are named as the reference names them, e.g.
``"Zernike Coefficients (Fourier (Chebyshev ()))"`` = Zernike of the Fourier
magnitude of the Chebyshev transform (innermost transform first,
src/FeatureNames.cpp:171-185).  Every other algorithm of the reference
(histograms, counts, Haralick, objects) has an integer-valued or piecewise
constant output: it can be added to a plan as a masked group, computed
exactly by the feature path and given no gradient.

Compound transforms are recomputed on the backward pass: a node's autograd
record keeps only the root planes, and the backward pass re-runs the
transform chain before applying the adjoints in reverse order.

All tensors are float64 on the GPU of the library's context (WNDGPU_DEVICE,
default 0); kernels run on torch's current stream.  There is no CPU
implementation: without the library or a GPU every call raises.
"""
import ctypes
import re

import numpy as np
import torch

from . import (WndGpuError, _Image, _check, _context, _load, fisher_weights, normalize_train,
               FeatureComputationPlan)

__all__ = ['FOURIER', 'CHEBYSHEV', 'ZERNIKE', 'CHEBYSHEV_COEFFS', 'CHEB_FOURIER', 'COMB', 'GABOR', 'op_forward',
           'op_backward', 'fourier', 'chebyshev', 'zernike', 'chebyshev_coeffs', 'cheb_fourier', 'comb_moments',
           'gabor_energy', 'DifferentiablePlan', 'WNDModel', 'parse_group']

FOURIER, CHEBYSHEV, ZERNIKE, CHEBYSHEV_COEFFS, CHEB_FOURIER, COMB, GABOR = 0, 1, 2, 3, 4, 5, 6
GABOR_MEANS = 100  # plan-level: the means of the 8 Gabor energy maps (op GABOR, averaged here)
TRANSFORMS = {'Fourier': FOURIER, 'Chebyshev': CHEBYSHEV}
ALGORITHMS = {'Zernike Coefficients': ZERNIKE, 'Chebyshev Coefficients (raw)': CHEBYSHEV_COEFFS,
              'Chebyshev-Fourier Coefficients (raw)': CHEB_FOURIER, 'Comb Moments (raw)': COMB,
              'Gabor Textures (raw)': GABOR_MEANS}
RAW_NAMES = ('Chebyshev Coefficients (raw)', 'Chebyshev-Fourier Coefficients (raw)', 'Comb Moments (raw)',
             'Gabor Textures (raw)')

_P = ctypes.c_void_p
_I = ctypes.c_int
_SIG = {
    'wndgpu_ad_output_shape': (_I, [_I, _I, _I, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    'wndgpu_ad_work_bytes': (ctypes.c_size_t, [_I, _I, _I, _I]),
    'wndgpu_ad_forward': (_I, [_P, _I, _P, _I, _I, _I, _P, _P, _P]),
    'wndgpu_ad_backward': (_I, [_P, _I, _P, _P, _I, _I, _I, _P, _P, _P]),
    'wndgpu_ad_wnd_work_bytes': (ctypes.c_size_t, [_I, _I, _I, _I]),
    'wndgpu_ad_wnd_forward': (_I, [_P, _P, _P, _I, _I, _I, _P, _P, _P, _P, _I, _P, _P, _P]),
    'wndgpu_ad_wnd_backward': (_I, [_P, _P, _P, _I, _I, _I, _P, _P, _P, _P, _I, _P, _P, _P, _P]),
}
_lib = None


def _ad_lib():
    global _lib
    if _lib is None:
        lib = _load()
        for name, (res, args) in _SIG.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def _vp(t):
    return _P(t.data_ptr())


def _stream(t):
    return _P(torch.cuda.current_stream(t.device).cuda_stream)


def _planes(x, what):
    if not isinstance(x, torch.Tensor) or not x.is_cuda or x.dtype != torch.float64 or x.dim() != 3:
        raise WndGpuError('%s: expected a float64 CUDA tensor [n, h, w]' % what)
    return x.contiguous()


def output_shape(op, h, w):
    oh, ow = _I(), _I()
    if _ad_lib().wndgpu_ad_output_shape(op, h, w, ctypes.byref(oh), ctypes.byref(ow)) != 0:
        raise WndGpuError('differentiable operation %d does not take %dx%d planes' % (op, w, h))
    return oh.value, ow.value


def _work(nbytes, like):
    return torch.empty(max(int(nbytes), 8), dtype=torch.uint8, device=like.device)


def op_forward(op, x):
    """Forward value of one operation on planes x [n, h, w]: [n, oh, ow]."""
    x = _planes(x, 'op_forward')
    n, h, w = x.shape
    oh, ow = output_shape(op, h, w)
    lib = _ad_lib()
    out = torch.empty((n, oh, ow), dtype=torch.float64, device=x.device)
    work = _work(lib.wndgpu_ad_work_bytes(op, n, h, w), x)
    _check(lib.wndgpu_ad_forward(_context(), op, _vp(x), n, h, w, _vp(out), _vp(work), _stream(x)), 'ad forward')
    return out


def op_backward(op, x, grad_out):
    """Adjoint of one operation at planes x: grad_out [n, oh, ow] -> [n, h, w]."""
    x = _planes(x, 'op_backward')
    n, h, w = x.shape
    oh, ow = output_shape(op, h, w)
    g = grad_out.reshape(n, oh, ow).to(torch.float64).contiguous()
    lib = _ad_lib()
    grad = torch.empty_like(x)
    work = _work(lib.wndgpu_ad_work_bytes(op, n, h, w), x)
    _check(lib.wndgpu_ad_backward(_context(), op, _vp(x), _vp(g), n, h, w, _vp(grad), _vp(work), _stream(x)),
           'ad backward')
    return grad


class _Op(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, op):
        ctx.save_for_backward(x)
        ctx.op = op
        return op_forward(op, x)

    @staticmethod
    def backward(ctx, g):
        x, = ctx.saved_tensors
        return op_backward(ctx.op, x, g), None


def fourier(x):
    """|DFT2| planes (the Fourier transform node), differentiable."""
    return _Op.apply(_planes(x, 'fourier'), FOURIER)


def chebyshev(x):
    """Chebyshev transform of order min(h, w), differentiable."""
    return _Op.apply(_planes(x, 'chebyshev'), CHEBYSHEV)


def zernike(x):
    """The 72 Zernike coefficients of each plane, [n, 72], differentiable."""
    return _Op.apply(_planes(x, 'zernike'), ZERNIKE).reshape(x.shape[0], -1)


def chebyshev_coeffs(x):
    """The order-20 Chebyshev coefficients, [n, 20, 20], differentiable."""
    return _Op.apply(_planes(x, 'chebyshev_coeffs'), CHEBYSHEV_COEFFS)


def cheb_fourier(x):
    """The NN x NN Chebyshev-Fourier coefficient magnitudes (after the
    Downsample of planes above 300 x 300), [n, NN*NN], differentiable."""
    return _Op.apply(_planes(x, 'cheb_fourier'), CHEB_FOURIER).reshape(x.shape[0], -1)


def comb_moments(x):
    """The 4 combs x (mean, std, skew, kurtosis) x 20 line moments, [n, 320],
    differentiable."""
    return _Op.apply(_planes(x, 'comb_moments'), COMB).reshape(x.shape[0], -1)


def gabor_energy(x):
    """The 8 Gabor energy maps (f0 = 0.1, then 1..7), [n, 8, h, w], differentiable."""
    n, h, w = x.shape
    return _Op.apply(_planes(x, 'gabor_energy'), GABOR).reshape(n, 8, h, w)


def _alg_forward(a, p):
    if a == GABOR_MEANS:
        return op_forward(GABOR, p).reshape(p.shape[0], 8, -1).mean(2)
    return op_forward(a, p).reshape(p.shape[0], -1)


def _alg_backward(a, p, g):
    if a == GABOR_MEANS:
        n, h, w = p.shape
        gE = (g.to(torch.float64) / (h * w)).reshape(n, 8, 1).expand(n, 8, h * w)
        return op_backward(GABOR, p, gE.reshape(n, 8 * h, w))
    return op_backward(a, p, g)


class _Node(torch.autograd.Function):
    """Features of the algorithms `algs` on the node chain(x); keeps only x."""

    @staticmethod
    def forward(ctx, x, chain, algs):
        p = x
        for t in chain:
            p = op_forward(t, p)
        outs = [_alg_forward(a, p) for a in algs]
        ctx.save_for_backward(x)
        ctx.chain, ctx.algs = chain, algs
        ctx.sizes = [o.shape[1] for o in outs]
        return torch.cat(outs, 1)

    @staticmethod
    def backward(ctx, g):
        x, = ctx.saved_tensors
        planes = [x]
        for t in ctx.chain:  # recompute the compound transform
            planes.append(op_forward(t, planes[-1]))
        top = planes[-1]
        gt, off = None, 0
        for a, k in zip(ctx.algs, ctx.sizes):
            ga = _alg_backward(a, top, g[:, off:off + k])
            gt = ga if gt is None else gt + ga
            off += k
        for t, p in zip(reversed(ctx.chain), reversed(planes[:-1])):
            gt = op_backward(t, p, gt)
        return gt, None, None


def parse_group(name):
    """'Alg (T1 (T2 ()))' -> ('Alg', [T2, T1]) : transform names in application order."""
    m = re.match(r'^\s*(.+?)\s*\((.*)\)\s*$', name)
    if not m:
        raise WndGpuError('feature group name "%s" has no transform part' % name)
    alg, inner = m.group(1), m.group(2).strip()
    outer = []
    while inner:
        m = re.match(r'^\s*(.+?)\s*\((.*)\)\s*$', inner)
        if not m:
            raise WndGpuError('cannot parse transform chain of "%s"' % name)
        outer.append(m.group(1))
        inner = m.group(2).strip()
    return alg, outer[::-1]


def _split_raw(name):
    # "Chebyshev Coefficients (raw) (Fourier ())": the algorithm name carries "(raw)"
    for raw in RAW_NAMES:
        if name.strip().startswith(raw):
            alg, chain = parse_group('X' + name.strip()[len(raw):])
            return raw, chain
    return parse_group(name)


class DifferentiablePlan(object):
    """Feature groups over FP64 device planes, differentiable where the
    feature is smooth.

    groups: reference-style names.  Algorithms in ALGORITHMS get gradients;
    any other algorithm of the feature path is computed by it exactly and
    masked (no gradient), which needs `masked=True` to be accepted.
    """

    def __init__(self, groups, masked=False):
        self.groups = list(groups)
        self.columns = []  # (start, size, group)
        self._nodes = {}   # chain tuple -> [algorithm ops]
        self._masked = []
        off = 0
        for g in self.groups:
            alg, chain = _split_raw(g)
            for t in chain:
                if t not in TRANSFORMS and alg in ALGORITHMS:
                    raise WndGpuError('transform "%s" in "%s" has no adjoint' % (t, g))
            if alg in ALGORITHMS:
                op = ALGORITHMS[alg]
                k = {ZERNIKE: 72, CHEBYSHEV_COEFFS: 400, CHEB_FOURIER: 529, COMB: 320, GABOR_MEANS: 8}[op]
                self._nodes.setdefault(tuple(TRANSFORMS[t] for t in chain), []).append((op, off, k))
            else:
                if not masked:
                    raise WndGpuError('"%s" is not differentiable (pass masked=True to compute it without a '
                                      'gradient)' % g)
                plan = FeatureComputationPlan('masked')
                plan.add(g)
                plan.finalize()
                k = plan.n_features
                self._masked.append((plan, off, k))
            self.columns.append((off, k, g))
            off += k
        self.n_features = off

    def differentiable_mask(self):
        m = np.zeros(self.n_features, bool)
        for ops in self._nodes.values():
            for _, off, k in ops:
                m[off:off + k] = True
        return m

    def __call__(self, x):
        """x: [n, h, w] float64 CUDA planes -> features [n, n_features]."""
        x = _planes(x, 'DifferentiablePlan')
        pieces = []
        for chain, ops in self._nodes.items():
            f = _Node.apply(x, chain, tuple(o for o, _, _ in ops))
            if f.shape[1] != sum(k for _, _, k in ops):
                # the raw coefficient groups have fixed sizes: 20 x 20 Chebyshev coefficients need a
                # node plane of at least 20 x 20, 23 x 23 Chebyshev-Fourier ones at least 23 x 23
                raise WndGpuError('node planes of these %dx%d images are too small for the raw coefficient groups'
                                  % (x.shape[2], x.shape[1]))
            c = 0
            for _, off, k in ops:
                pieces.append((off, f[:, c:c + k]))
                c += k
        if self._masked:
            xd = x.detach().contiguous()
            torch.cuda.current_stream(x.device).synchronize()
            for plan, off, k in self._masked:
                pieces.append((off, _compute_device(plan, xd)))
        pieces.sort(key=lambda p: p[0])
        return torch.cat([p[1] for p in pieces], 1)


def _compute_device(plan, x):
    """The feature path on device planes x [n, h, w] (float64, synchronised):
    [n, plan.n_features], no gradient."""
    n, h, w = x.shape
    structs = (_Image * n)()
    for i in range(n):
        structs[i] = _Image(_P(x[i].data_ptr()), h, w, w, 2)
    out = torch.empty((n, plan.n_features), dtype=torch.float64, device=x.device)
    _check(_load().wndgpu_compute_device(_context(), plan.handle, structs, n, _vp(out), None), 'wndgpu_compute_device')
    return out


class _WND(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, model):
        ctx.save_for_backward(x)
        ctx.model = model
        return model._forward(x)

    @staticmethod
    def backward(ctx, g):
        x, = ctx.saved_tensors
        return ctx.model._backward(x, g.to(torch.float64).contiguous()), None


class WNDModel(object):
    """A WND5 training set on the device: normalised training matrix, its
    per-feature min / max, Fisher weights (TrainingSet::SetFisherScores with
    the top `used_fraction` kept) and classes 1..n_classes.  `probabilities(x)`
    takes raw feature vectors [n, F] and is differentiable."""

    def __init__(self, train_features, classes, n_classes, used_fraction=0.15, device=0):
        raw = np.ascontiguousarray(np.asarray(train_features, np.float64))
        cls = np.ascontiguousarray(np.asarray(classes, np.int32))
        if raw.ndim != 2 or cls.shape != (raw.shape[0],):
            raise WndGpuError('WND5 training set: %s feature rows and %s class labels' % (raw.shape[:1], cls.shape))
        if n_classes < 1 or n_classes > 63 or cls.min() < 0 or cls.max() > n_classes:
            raise WndGpuError('WND5 training set: class labels must lie in 0..n_classes (n_classes <= 63)')
        norm, mins, maxs = normalize_train(raw)
        w = fisher_weights(norm, cls, n_classes, used_fraction)
        dev = torch.device('cuda', device)
        self.n_classes = int(n_classes)
        self.n_samples, self.n_features = norm.shape
        self.train = torch.from_numpy(norm).to(dev)
        self.classes = torch.from_numpy(cls).to(dev)
        self.weights = torch.from_numpy(w).to(dev)
        self.mins = torch.from_numpy(mins).to(dev)
        self.maxs = torch.from_numpy(maxs).to(dev)

    def _args(self):
        return (_vp(self.train), _vp(self.classes), self.n_samples, self.n_features, self.n_classes,
                _vp(self.weights), _vp(self.mins), _vp(self.maxs))

    def _check_x(self, x):
        if not x.is_cuda or x.dtype != torch.float64 or x.dim() != 2 or x.shape[1] != self.n_features:
            raise WndGpuError('WND5: expected float64 CUDA feature vectors [n, %d]' % self.n_features)
        return x.contiguous()

    def _forward(self, x):
        x = self._check_x(x)
        lib = _ad_lib()
        n = x.shape[0]
        prob = torch.empty((n, self.n_classes), dtype=torch.float64, device=x.device)
        work = _work(lib.wndgpu_ad_wnd_work_bytes(self.n_samples, self.n_features, self.n_classes, n), x)
        _check(lib.wndgpu_ad_wnd_forward(_context(), *self._args(), _vp(x), n, _vp(prob), _vp(work), _stream(x)),
               'ad WND forward')
        return prob

    def _backward(self, x, g):
        x = self._check_x(x)
        lib = _ad_lib()
        n = x.shape[0]
        gx = torch.empty_like(x)
        work = _work(lib.wndgpu_ad_wnd_work_bytes(self.n_samples, self.n_features, self.n_classes, n), x)
        _check(lib.wndgpu_ad_wnd_backward(_context(), *self._args(), _vp(x), n, _vp(g), _vp(gx), _vp(work),
                                          _stream(x)), 'ad WND backward')
        return gx

    def probabilities(self, x):
        """WND5 marginal probabilities [n, n_classes] of raw feature vectors x."""
        return _WND.apply(x, self)
