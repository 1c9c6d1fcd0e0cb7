"""Feature-scaling utilities of the wndcharm package (wndcharm/utils.py in the
reference), computed on the GPU through the C ABI (wndgpu_py_normalize).

Same names, arguments, in-place behaviour and return values as the
reference's ReplaceNonReal (utils.py:124-169) and normalize_by_columns
(utils.py:172-240)."""
import ctypes

import numpy as np

from . import WndGpuError, _check, _context, _load, _ptr

ZSCORE, NON_REAL_CHECK, GIVEN_MINS, GIVEN_MAXS, GIVEN_MEANS, GIVEN_STDEVS, REPLACE_ONLY = 1, 2, 4, 8, 16, 32, 64


def _run(feature_matrix, flags, mins, maxs, means, stdevs):
    if not isinstance(feature_matrix, np.ndarray) or feature_matrix.dtype != np.float64:
        raise WndGpuError('expected a float64 numpy array')
    one_d = feature_matrix.ndim == 1
    M = np.ascontiguousarray(feature_matrix.reshape(1, -1) if one_d else feature_matrix)
    if M.ndim != 2:
        raise WndGpuError('expected a 1-D or 2-D feature matrix')
    n, f = M.shape
    arrs = []
    for a, bit in ((mins, GIVEN_MINS), (maxs, GIVEN_MAXS), (means, GIVEN_MEANS), (stdevs, GIVEN_STDEVS)):
        if a is not None:
            a = np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1)).copy()
            if a.shape != (f,):
                raise WndGpuError('reference values of %d features for a matrix of %d' % (a.size, f))
            flags |= bit
        else:
            a = np.zeros(f)
        arrs.append(a)
    replaced = ctypes.c_int()
    _check(_load().wndgpu_py_normalize(_context(), _ptr(M), n, f, flags, *[_ptr(a) for a in arrs],
                                       ctypes.byref(replaced)), 'normalize_by_columns')
    if M is not feature_matrix and not (one_d and np.shares_memory(M, feature_matrix)):
        feature_matrix[...] = M.reshape(feature_matrix.shape)
    return bool(replaced.value), arrs


def ReplaceNonReal(feature_matrix, mins=None, maxs=None):
    """In place: NaN -> 0 and +/-inf clipped (utils.py:124-169, including its
    rebinding of the bounds when only some are given).  Returns True if the
    matrix held a NaN or +/-inf."""
    replaced, _ = _run(feature_matrix, REPLACE_ONLY, mins, maxs, None, None)
    return replaced


def normalize_by_columns(feature_matrix, mins=None, maxs=None, means=None, stdevs=None, zscore=False,
                         non_real_check=True):
    """In place: min-max scaling to 0..100 (or z-scores), zero-range columns
    set to 0 (utils.py:172-240).  Returns (mins, maxs, means, stdevs): the
    values given, or those derived here (None for the other form's pair)."""
    flags = (ZSCORE if zscore else 0) | (NON_REAL_CHECK if non_real_check else 0)
    _, (mn, mx, me, sd) = _run(feature_matrix, flags, mins, maxs, means, stdevs)
    if not zscore:
        return (mn if mins is None else mins, mx if maxs is None else maxs, means, stdevs)
    return (mins, maxs, me if means is None else means, sd if stdevs is None else stdevs)
