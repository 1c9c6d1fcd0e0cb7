"""Shared test helpers: oracle / reference-harness loaders and the
reference's own feature comparator."""
import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, 'tests', 'golden')
ORACLE_SO = os.path.join(ROOT, 'oracle', 'liboracle.so')
REF_SO = os.path.join(ROOT, 'oracle', '_ref', 'libwndref.so')

P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            pytest.skip('oracle not built (run __graft_entry__.build())')
        lib = ctypes.CDLL(ORACLE_SO)
        lib.orc_transform.argtypes = [I, P, I, I, P, P, P]
        lib.orc_algorithm.argtypes = [I, P, I, I, P]
        lib.orc_features.argtypes = [I, P, I, I, P]
        lib.orc_feature_name.restype = ctypes.c_char_p
        lib.orc_otsu.argtypes = [P, I, I]
        lib.orc_cf_magnitudes.argtypes = [P, I, I, P, P]
        lib.orc_comb_raw.argtypes = [P, I, I, P]
        lib.orc_gabor_energy.argtypes = [P, I, I, P]
        lib.orc_otsu.restype = D
        lib.orc_haralick_grays.argtypes = [P, I, I, P]
        lib.orc_radon_proj.argtypes = [P, I, I, P, P]
        lib.orc_cooc.argtypes = [P, I, I, I, P, P]
        lib.orc_bwlabel.argtypes = [P, I, I, P]
        lib.orc_bwlabel.restype = ctypes.c_ulong
        lib.orc_haralick_f14.argtypes = [P, I, P, P]
        lib.orc_normalize_train.argtypes = [P, I, I, P, P]
        lib.orc_normalize_test.argtypes = [P, I, P, P]
        lib.orc_fisher.argtypes = [P, P, I, I, I, D, P]
        lib.orc_wnd5.argtypes = [P, P, I, I, I, P, P, P, P]
        lib.orc_wnd5.restype = ctypes.c_long
        lib.orc_wnn.argtypes = [P, P, I, I, I, P, P, P, P]
        lib.orc_wnn.restype = ctypes.c_long
        lib.orc_plan_n_features.argtypes = [I]
        lib.orc_features_rgb.argtypes = [I, P, I, I, I, P]
        lib.orc_rgb_planes.argtypes = [P, I, I, I, P, P]
        lib.orc_rgb_to_hsv.argtypes = [I, I, I, P]
        lib.orc_find_color.restype = ctypes.c_long
        lib.orc_find_color.argtypes = [I, I, I]
        lib.orc_find_color_many.argtypes = [P, ctypes.c_long, P]
        lib.orc_color_transform.argtypes = [I, P, I, I, I, P]
        lib.orc_color_histogram.argtypes = [P, I, I, I, P]
        _oracle = lib
    return _oracle


def ref_harness():
    if not os.path.exists(REF_SO):
        pytest.skip('reference harness oracle/_ref not built (needs /root/reference)')
    lib = ctypes.CDLL(REF_SO)
    lib.ref_texture.argtypes = [P, I, I, I, P]
    lib.ref_eigen.argtypes = [P, I, P, P]
    lib.ref_radon.argtypes = [P, I, I, I, I, P]
    lib.ref_wavelet.argtypes = [P, I, I, P, P, P]
    lib.ref_cheb1d.argtypes = [P, I, I, I, P]
    lib.ref_find_color_many.argtypes = [P, ctypes.c_long, P]
    return lib


ALG_NFEAT = [32, 32, 72, 28, 24, 6, 48, 12, 20, 5, 28, 34, 34, 7, 1, 20]
T_FOURIER, T_CHEBYSHEV, T_WAVELET, T_EDGE, T_COLOR, T_HUE, T_COLOR_TRANSFORM = 0, 1, 2, 3, 4, 5, 6
# feature counts of the standard sets: short, long, short colour, long colour
# (src/wndchrm_src/signatures.h:55-57 for the colour sets)
SET_NFEAT = [1059, 2919, 2199, 4059]


def _set_of(long_set):
    return int(long_set) if isinstance(long_set, bool) else long_set


def orc_features(img, long_set=True):
    """oracle features of a gray plane; long_set: True/False or a set 0..1"""
    img = np.ascontiguousarray(img, dtype=np.float64)
    lib = oracle()
    s = _set_of(long_set)
    out = np.zeros(SET_NFEAT[s])
    assert lib.orc_features(s, ptr(img), img.shape[0], img.shape[1], ptr(out)) == 0
    return out


def orc_features_rgb(rgb, set_):
    """oracle features of an (h, w, 3) uint8 / uint16 RGB image, sets 0..3"""
    rgb = np.ascontiguousarray(rgb)
    bits = 8 if rgb.dtype == np.uint8 else 16
    out = np.zeros(SET_NFEAT[set_])
    assert oracle().orc_features_rgb(set_, ptr(rgb), bits, rgb.shape[0], rgb.shape[1], ptr(out)) == 0
    return out


def orc_rgb_planes(rgb):
    """(gray plane, (h, w, 3) HSV bytes) of an RGB image"""
    rgb = np.ascontiguousarray(rgb)
    h, w = rgb.shape[:2]
    gray = np.zeros((h, w))
    hsv = np.zeros((h, w, 3), np.uint8)
    assert oracle().orc_rgb_planes(ptr(rgb), 8 if rgb.dtype == np.uint8 else 16, h, w, ptr(gray), ptr(hsv)) == 0
    return gray, hsv


def orc_color_transform(kind, rgb):
    rgb = np.ascontiguousarray(rgb)
    h, w = rgb.shape[:2]
    out = np.zeros((h, w))
    assert oracle().orc_color_transform(kind, ptr(rgb), 8 if rgb.dtype == np.uint8 else 16, h, w, ptr(out)) == 0
    return out


def orc_color_histogram(rgb):
    rgb = np.ascontiguousarray(rgb)
    out = np.zeros(20)
    assert oracle().orc_color_histogram(ptr(rgb), 8 if rgb.dtype == np.uint8 else 16, rgb.shape[0], rgb.shape[1],
                                        ptr(out)) == 0
    return out


def orc_find_color_many(hsv):
    hsv = np.ascontiguousarray(hsv, dtype=np.uint8).reshape(-1, 3)
    out = np.zeros(len(hsv), np.int64)
    oracle().orc_find_color_many(ptr(hsv), len(hsv), ptr(out))
    return out


class OrcSampling(ctypes.Structure):
    _fields_ = [('roi_x', I), ('roi_y', I), ('roi_w', I), ('roi_h', I), ('downsample', I), ('mean', D),
                ('stdev', D), ('rotations', I), ('tiles_x', I), ('tiles_y', I)]


def orc_features_sampled(img, set_, roi=None, downsample=0, mean=0.0, stdev=0.0, rotations=1, tiles=(1, 1)):
    """oracle features [n_samples][n_features] of an image (2-D uint8 / uint16 /
    float64, or (h, w, 3) uint8 / uint16) under the sampling options"""
    img = np.ascontiguousarray(img)
    if img.ndim == 3:
        dt = 4 if img.dtype == np.uint8 else 5
    else:
        dt = {np.dtype(np.uint16): 0, np.dtype(np.float64): 2, np.dtype(np.uint8): 3}[img.dtype]
    x, y, w, h = roi if roi is not None else (-1, -1, 0, 0)
    s = OrcSampling(x, y, w, h, downsample, mean, stdev, rotations, tiles[0], tiles[1])
    lib = oracle()
    lib.orc_features_sampled.argtypes = [I, P, I, I, I, ctypes.POINTER(OrcSampling), P]
    ns = rotations * tiles[0] * tiles[1]
    out = np.zeros((ns, SET_NFEAT[set_]))
    rc = lib.orc_features_sampled(set_, ptr(img), dt, img.shape[0], img.shape[1], ctypes.byref(s), ptr(out))
    return out if rc == 0 else None


def orc_names(long_set=True):
    lib = oracle()
    s = _set_of(long_set)
    return [lib.orc_feature_name(s, i).decode() for i in range(SET_NFEAT[s])]


def orc_transform(kind, img):
    img = np.ascontiguousarray(img, dtype=np.float64)
    h, w = img.shape
    out = np.zeros((2 * ((h + 9) // 2) + 2) * (2 * ((w + 9) // 2) + 2))
    oh, ow = ctypes.c_int(), ctypes.c_int()
    assert oracle().orc_transform(kind, ptr(img), h, w, ptr(out), ctypes.byref(oh), ctypes.byref(ow)) == 0
    return out[:oh.value * ow.value].reshape(oh.value, ow.value)


def orc_algorithm(alg, img):
    img = np.ascontiguousarray(img, dtype=np.float64)
    out = np.zeros(ALG_NFEAT[alg])
    assert oracle().orc_algorithm(alg, ptr(img), img.shape[0], img.shape[1], ptr(out)) == 0
    return out


def wndcharm_compare(a, b):
    """The reference's feature comparator (wndcharm/utils.py:276-366): equal,
    or |a-b| < 1e-7, or the 6-significant-digit mantissas differ by at most
    one unit in the last digit."""
    if a == b or abs(a - b) < 1e-7:
        return True
    if not (np.isfinite(a) and np.isfinite(b)):
        return False
    a6, b6 = float('%.6g' % a), float('%.6g' % b)
    mag = max(abs(a6), abs(b6))
    if mag == 0:
        return True
    unit = 10.0 ** (np.floor(np.log10(mag)) - 5)
    return abs(a6 - b6) <= unit * 1.0001


def load_npz(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


# Feature indices whose reference values are integer counts or ratios of
# integer counts (bit-exact bar); the rest are floating point.
ALG_NAMES = ['Chebyshev-Fourier Coefficients', 'Chebyshev Coefficients', 'Zernike Coefficients', 'Haralick Textures',
             'Multiscale Histograms', 'Tamura Textures', 'Comb Moments', 'Radon Coefficients', 'Fractal Features',
             'Pixel Intensity Statistics', 'Edge Features', 'Otsu Object Features', 'Inverse-Otsu Object Features',
             'Gabor Textures', 'Gini Coefficient', 'Color Histogram']
EXACT = {
    0: list(range(32)), 1: list(range(32)), 2: [], 3: [], 4: list(range(24)), 5: [0, 1, 2, 5],
    6: list(range(48)), 7: list(range(12)), 8: [], 9: [1, 3, 4],
    10: [0] + list(range(1, 14)) + list(range(17, 25)) + [26],
    11: list(range(0, 14)) + [15, 16, 17] + list(range(18, 29)) + [30, 31, 33],
    12: list(range(0, 14)) + [15, 16, 17] + list(range(18, 29)) + [30, 31, 33],
    13: list(range(7)), 14: [], 15: list(range(20)),
}


def exact_mask_for_names(names):
    """bool mask over plan columns: True where the feature is integer-valued."""
    idx = {n: i for i, n in enumerate(ALG_NAMES)}
    m = np.zeros(len(names), bool)
    for c, nm in enumerate(names):
        g, k = nm.rsplit(' [', 1)
        k = int(k[:-1])
        alg = g[:g.index(' (')]
        m[c] = k in EXACT[idx[alg]]
    return m


def features_in_context(env, imgs, plan):
    """wndcharm.compute_features(imgs, plan) in a fresh engine context created
    with the environment variables env set (the engine reads its switches at
    creation); the context is destroyed afterwards"""
    import wndcharm
    lib = wndcharm._load()
    old = wndcharm._context()
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    c = ctypes.c_void_p()
    try:
        rc = lib.wndgpu_create(0, ctypes.byref(c))
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k)
            else:
                os.environ[k] = v
    assert rc == 0
    wndcharm._ctx = c
    try:
        return wndcharm.compute_features(imgs, plan)
    finally:
        wndcharm._ctx = old
        lib.wndgpu_destroy(c)
