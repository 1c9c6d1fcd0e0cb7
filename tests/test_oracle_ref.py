"""Pins oracle functions against the reference's own code compiled from its
sources (oracle/ref/Makefile -> oracle/_ref/libwndref.so):
  * Haralick f1..f14 per angle: src/textures/haralick/CVIPtexture.cpp:384-1023,
    f14 through the vendored Eigen 3.1.2 EigenSolver (eigenvalue index 2 in
    Schur-diagonal order);
  * Symlet-5 wavelet transform: src/transforms/wavelet/ (Symlet5, Wavelet::transform2D);
  * Radon projections: src/transforms/radon.cpp:14-91;
  * Chebyshev transform contractions: src/transforms/chebyshev.cpp:38-116.
Skipped where /root/reference was not available to build the harness."""
import ctypes

import numpy as np
import pytest

from helpers import (ALG_NFEAT, T_CHEBYSHEV, T_FOURIER, T_WAVELET, load_npz, oracle, orc_algorithm,
                     orc_transform, ptr, ref_harness)


def golden_image(name='img_t1_s01_c05_ij'):
    return load_npz(name + '.npz')['pixels'].astype(np.float64)


def synth(h, w, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    img = 500 + 200 * np.sin(2 * np.pi * (xx * 0.3 + yy * 0.8) / 9.0) + rng.poisson(50, (h, w))
    return np.clip(np.round(img), 0, 4095).astype(np.float64)


def grays_of(img):
    h, w = img.shape
    g = np.zeros(h * w, np.uint8)
    src = np.ascontiguousarray(img, np.float64)  # held by name across the call
    oracle().orc_haralick_grays(ptr(src), h, w, ptr(g))
    return g


def haralick_from_ref(img):
    lib = ref_harness()
    h, w = img.shape
    g = grays_of(img)
    per = []
    for ang in (0, 45, 90, 135):
        out = np.zeros(14)
        assert lib.ref_texture(ptr(g), h, w, ang, ptr(out)) == 0
        per.append(out)
    per = np.array(per)
    s = np.zeros(14)
    for a in range(4):
        s = s + per[a]
    temp = np.concatenate([s / 4, per.max(0) - per.min(0)])
    order = [0, 14, 1, 15, 2, 16, 9, 23, 10, 24, 8, 22, 11, 25, 4, 18, 13, 27, 12, 26, 5, 19, 7, 21, 6, 20, 3, 17]
    return temp[order]


@pytest.mark.parametrize('src', ['root', 'fourier', 'wavelet', 'synth'])
def test_haralick_vs_reference(src):
    img = golden_image()
    if src == 'fourier':
        img = orc_transform(T_FOURIER, img)
    elif src == 'wavelet':
        img = orc_transform(T_WAVELET, img)
    elif src == 'synth':
        img = synth(96, 80, 3)
    want = haralick_from_ref(img)
    got = orc_algorithm(3, img)
    # f1..f13 use plain loops in the reference; f14 (Q, Hessenberg, Francis QR)
    # follows Eigen's kernel orders on row-major arrays: all 28 bit-exact
    assert np.array_equal(got, want), np.abs(got - want).max()


def test_f14_eigen_order():
    """The oracle's Hessenberg + Francis QR equals Eigen's EigenSolver on a
    row-major matrix (the reference's Q.matrix().eigenvalues()) bit for bit,
    eigenvalue order included, on Q-like matrices (row-stochastic,
    rank-deficient) of odd and even sizes."""
    lib = ref_harness()
    rng = np.random.default_rng(5)
    for n in (3, 5, 17, 64, 150, 201):
        A = rng.random((n, n)) ** 3
        A = A + A.T
        P = A / A.sum()
        px = P.sum(1)
        py = P.sum(0)
        Q = np.array([[(P[i] * P[j] / (px[i] * py + 1e-13)).sum() for j in range(n)] for i in range(n)])
        Q = np.ascontiguousarray(Q)
        r1, i1, r2, i2 = (np.zeros(n) for _ in range(4))
        lib.ref_eigen(ptr(Q), n, ptr(r1), ptr(i1))
        oracle().orc_haralick_f14(ptr(Q), n, ptr(r2), ptr(i2))
        assert np.array_equal(r1, r2) and np.array_equal(i1, i2), (n, np.abs(r1 - r2).max())


@pytest.mark.parametrize('shape', [(113, 113), (64, 90), (37, 20)])
def test_wavelet_vs_reference_bitexact(shape):
    img = synth(*shape, seed=shape[0])
    lib = ref_harness()
    h, w = shape
    out = np.zeros((h + 20) * (w + 20))
    oh, ow = ctypes.c_int(), ctypes.c_int()
    lib.ref_wavelet(ptr(img), h, w, ptr(out), ctypes.byref(oh), ctypes.byref(ow))
    want = out[:oh.value * ow.value].reshape(oh.value, ow.value)
    got = orc_transform(T_WAVELET, img)
    assert got.shape == want.shape == (2 * ((h + 9) // 2), 2 * ((w + 9) // 2))
    assert np.array_equal(got, want)


@pytest.mark.parametrize('shape', [(113, 113), (64, 90), (90, 64)])
def test_radon_vs_reference_bitexact(shape):
    img = synth(*shape, seed=7)
    h, w = shape
    mine = np.zeros(4 * (2 * (h + w) + 8))
    rs = ctypes.c_int()
    oracle().orc_radon_proj(ptr(img), h, w, ptr(mine), ctypes.byref(rs))
    r = rs.value
    colmajor = np.ascontiguousarray(img.T).ravel()
    ref = np.zeros(4 * r)
    ref_harness().ref_radon(ptr(colmajor), h, w, -((r - 1) // 2), r, ptr(ref))
    assert np.array_equal(mine[:4 * r], ref)


@pytest.mark.parametrize('shape', [(64, 64), (50, 70), (70, 50)])
def test_chebyshev_vs_reference_bitexact(shape):
    img = synth(*shape, seed=11)
    h, w = shape
    n = min(h, w)
    lib = ref_harness()
    a = np.zeros(h * n)
    lib.ref_cheb1d(ptr(img), h, w, n, ptr(a))
    t = np.ascontiguousarray(a.reshape(h, n).T)
    b = np.zeros(n * n)
    lib.ref_cheb1d(ptr(t), n, h, n, ptr(b))
    got = orc_transform(T_CHEBYSHEV, img)
    assert np.array_equal(got, b.reshape(n, n))


@pytest.mark.parametrize('shape', [(128, 128), (113, 113), (70, 100), (61, 122), (33, 97)])
def test_fourier_is_the_dft(shape):
    """The oracle's Fourier node (ImageMatrix::fft2, cmatrix.cpp:1107-1160) is
    |DFT2| of the plane with the reference's Hermitian fill of rows >= H/2+1;
    checked against numpy's FFT (an independent implementation)."""
    img = synth(*shape, seed=3)
    got = orc_transform(T_FOURIER, img)
    mag = np.abs(np.fft.fft2(img))
    h, w = shape
    hh = h // 2 + 1
    np.testing.assert_allclose(got[:hh], mag[:hh], rtol=1e-10, atol=1e-9 * mag.max())
    # lower rows are exact copies of the mirrored upper rows
    for y in range(hh, h):
        assert got[y, 0] == got[h - y, 0]
        assert np.array_equal(got[y, 1:], got[h - y, :0:-1])
