"""Pins the oracle's sampling of an image before the features (OpenImage's
crop / Downsample / normalize, cmatrix.cpp:252-294, 579-720, 836-888;
Rotate 727-767; the tiles of TrainingSet::AddImageFile, TrainingSet.cpp:
1455-1480, in setup_featureset's order, wndchrm.cpp:89-128):
  * the reference's own pre-processing test (test_FeatureVector.py:222-296):
    t1_s01_c05_ij normalised to mean 1000 (and stdev 100) has that mean
    (and stdev), and a 25% downsample of it is int(0.25 * w) x int(0.25 * h);
  * the reference's tiling test (test_FeatureVector.py:152-217): tile (5, 4)
    of a 6 x 5 tiling of a 1388 x 1040 image is the 231 x 208 block at
    (1155, 832);
  * rotations against numpy: 90 degrees is np.rot90(k=-1), 180 is
    np.rot90(k=2), and the reference's "270" is a transpose;
  * tiles of rotated images with the reference's tile sizes; sample order;
  * Downsample of constant and linear images (the gray Downsample itself is
    pinned by the Chebyshev-Fourier golden of the 301 x 300 image);
  * the crop of an RGB16 image scales by the whole image's channel range."""
import ctypes

import numpy as np
import pytest

from helpers import P, ptr, load_npz, oracle, orc_rgb_planes


class Sampling(ctypes.Structure):
    _fields_ = [('roi_x', ctypes.c_int), ('roi_y', ctypes.c_int), ('roi_w', ctypes.c_int), ('roi_h', ctypes.c_int),
                ('downsample', ctypes.c_int), ('mean', ctypes.c_double), ('stdev', ctypes.c_double),
                ('rotations', ctypes.c_int), ('tiles_x', ctypes.c_int), ('tiles_y', ctypes.c_int)]


def sampling(roi=None, downsample=0, mean=0.0, stdev=0.0, rotations=1, tiles=(1, 1)):
    x, y, w, h = roi if roi is not None else (-1, -1, 0, 0)
    return Sampling(x, y, w, h, downsample, mean, stdev, rotations, tiles[0], tiles[1])


DT = {np.dtype(np.uint16): 0, np.dtype(np.float64): 2, np.dtype(np.uint8): 3}


def sample_plane(img, s, k, with_hsv=False):
    lib = oracle()
    lib.orc_sample_plane.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Sampling),
                                     ctypes.c_int, P, P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    img = np.ascontiguousarray(img)
    dt = DT[img.dtype] if img.ndim == 2 else (4 if img.dtype == np.uint8 else 5)
    h, w = img.shape[:2]
    cap = h * w
    plane = np.zeros(cap)
    hsv = np.zeros(cap * 3, np.uint8)
    oh, ow = ctypes.c_int(), ctypes.c_int()
    rc = lib.orc_sample_plane(ptr(img), dt, h, w, ctypes.byref(s), k, ptr(plane), ptr(hsv), ctypes.byref(oh),
                              ctypes.byref(ow))
    if rc != 0:
        return None
    n = oh.value * ow.value
    p = plane[:n].reshape(oh.value, ow.value)
    return (p, hsv[:3 * n].reshape(oh.value, ow.value, 3)) if with_hsv else p


def t1():
    return load_npz('img_t1_s01_c05_ij.npz')['pixels'].astype(np.uint16)


def test_reference_normalisation_test():
    img = t1()
    m = sample_plane(img, sampling(mean=1000), 0)
    assert int(m.mean()) == 1000
    ms = sample_plane(img, sampling(mean=1000, stdev=100), 0)
    assert int(round(ms.mean())) == 1000 and int(round(ms.std())) == 100
    # no clamping here, so the shift is exact up to rounding
    assert np.allclose(m - img, 1000 - img.mean(), atol=1e-9)


def test_reference_downsample_dimensions():
    img = t1()
    d = sample_plane(img, sampling(downsample=25), 0)
    assert d.shape == (int(113 * 0.25), int(113 * 0.25))


def test_reference_tiling_geometry():
    img = np.random.default_rng(1).integers(0, 256, (1040, 1388), dtype=np.uint8)
    s = sampling(tiles=(6, 5))
    k = 5 * 5 + 4  # tile x 5, tile y 4: tile y varies fastest
    assert np.array_equal(sample_plane(img, s, k), img[832:1040, 1155:1386])
    assert np.array_equal(sample_plane(img, s, 0), img[0:208, 0:231])


@pytest.mark.parametrize('shape', [(40, 56), (37, 37)])
def test_rotations(shape):
    img = np.random.default_rng(2).integers(0, 4096, shape).astype(np.uint16)
    s = sampling(rotations=4)
    assert np.array_equal(sample_plane(img, s, 0), img)
    assert np.array_equal(sample_plane(img, s, 1), np.rot90(img, -1))
    assert np.array_equal(sample_plane(img, s, 2), np.rot90(img, 2))
    assert np.array_equal(sample_plane(img, s, 3), img.T)


def test_tiles_of_rotations():
    for shape in ((60, 84), (72, 72)):
        img = np.random.default_rng(3).integers(0, 4096, shape).astype(np.uint16)
        s = sampling(rotations=2, tiles=(3, 3))
        for k in range(18):
            r, tx, ty = k // 9, (k % 9) // 3, k % 3
            rot = np.rot90(img, -r) if r else img
            rh, rw = rot.shape
            # rotations 1 and 3 take the tile height from the rotated width
            # and the tile width from the rotated height (TrainingSet.cpp:1468-1470)
            th, tw = (rw // 3, rh // 3) if r == 1 else (rh // 3, rw // 3)
            if (tx + 1) * tw > rw or (ty + 1) * th > rh:
                # past the rotated image: the reference's submatrix refuses it
                # (its tile stays the previous one); only square images avoid this
                assert sample_plane(img, s, k) is None, k
                continue
            want = rot[ty * th:(ty + 1) * th, tx * tw:(tx + 1) * tw]
            assert np.array_equal(sample_plane(img, s, k), want), k


def test_downsample_constant_and_ramp():
    c = np.full((50, 70), 321, np.uint16)
    d = sample_plane(c, sampling(downsample=40), 0)
    assert d.shape == (20, 28)
    assert np.allclose(d[:-1, :-1], 321, rtol=0, atol=1e-9)
    yy, xx = np.mgrid[0:50, 0:70]
    ramp = (3 * xx + 5 * yy).astype(np.uint16)
    d = sample_plane(ramp, sampling(downsample=50), 0)
    # each output averages a 2x2 block: the block centre of the ramp
    want = 3 * (2 * np.arange(35) + 0.5)[None, :] + 5 * (2 * np.arange(25) + 0.5)[:, None]
    assert np.allclose(d, want, atol=1e-9)


def test_crop_and_colour_planes():
    rng = np.random.default_rng(5)
    rgb = rng.integers(100, 60000, (48, 64, 3)).astype(np.uint16)
    gray, hsv = orc_rgb_planes(rgb)
    p, h = sample_plane(rgb, sampling(roi=(10, 6, 30, 20)), 0, with_hsv=True)
    assert np.array_equal(p, gray[6:26, 10:40]) and np.array_equal(h, hsv[6:26, 10:40])
    # the colour plane rotates and tiles with the pixels
    p, h = sample_plane(rgb, sampling(rotations=2), 1, with_hsv=True)
    assert np.array_equal(p, np.rot90(gray, -1)) and np.array_equal(h, np.rot90(hsv, -1))
    assert sample_plane(rgb, sampling(roi=(40, 0, 30, 20)), 0) is None  # outside the image


def test_colour_downsample_bytes():
    rgb = np.zeros((40, 40, 3), np.uint8)
    rgb[..., 0] = 200
    rgb[..., 2] = 40
    _, hsv = orc_rgb_planes(rgb)
    _, h = sample_plane(rgb, sampling(downsample=30), 0, with_hsv=True)
    # a constant's weighted sum divided back is the constant, or from
    # rounding just below it: each of the two passes truncates to a byte
    for ch in range(3):
        v = int(hsv[0, 0, ch])
        assert set(np.unique(h[:-1, :-1, ch]).tolist()) <= {v - 2, v - 1, v}
        assert v in set(np.unique(h[:-1, :-1, ch]).tolist())
