"""Pins the oracle's colour path:
  * FindColor (the fuzzy colour class of an HSV byte triple) against the
    reference's own colors/FuzzyCalc.cpp compiled from its sources
    (oracle/_ref), over a grid and over the triples of random RGB pixels;
  * RGB2HSV (src/cmatrix.h:95-126) against its closed forms on the primary
    and secondary colours and greys, and against Python's colorsys on random
    pixels (hue on a 240 scale, S and V on a 240 scale);
  * RGB2GRAY (src/cmatrix.h:158-160) on the unit colours;
  * the 16-bit RGB scaling (min / max over the channels) on an image whose
    range is exactly 0..255, where it must be the identity;
  * the colour plans' sizes (src/wndchrm_src/signatures.h:55-57) and layout:
    the short colour set is the short set followed by the colour groups, the
    long colour set holds the long set's groups in order.
The reference's test data hold no colour images or colour signatures, so
the colour features themselves are pinned through these functions."""
import colorsys

import numpy as np
import pytest

from helpers import (SET_NFEAT, T_COLOR, T_HUE, oracle, orc_color_histogram, orc_color_transform, orc_find_color_many,
                     orc_names, orc_rgb_planes, ptr, ref_harness)


def hsv_of(r, g, b):
    out = np.zeros(3, np.uint8)
    oracle().orc_rgb_to_hsv(r, g, b, ptr(out))
    return tuple(int(x) for x in out)


def test_find_color_vs_reference():
    lib = ref_harness()
    h, s, v = np.meshgrid(np.arange(256), np.arange(0, 256, 8), np.arange(0, 256, 8), indexing='ij')
    grid = np.stack([h, s, v], -1).reshape(-1, 3)
    rgb = np.random.default_rng(5).integers(0, 256, (256, 256, 3), dtype=np.uint8)
    _, hsv = orc_rgb_planes(rgb)
    tri = np.ascontiguousarray(np.concatenate([grid, hsv.reshape(-1, 3)]).astype(np.uint8))
    want = np.zeros(len(tri), np.int64)
    lib.ref_find_color_many(ptr(tri), len(tri), ptr(want))
    got = orc_find_color_many(tri)
    assert np.array_equal(got, want), np.flatnonzero(got != want)[:10]
    # every class 1..19 occurs, and nothing else
    assert set(np.unique(got)) == set(range(1, 20))


def test_rgb_to_hsv_closed_forms():
    # primaries and secondaries: hue at multiples of 60 degrees on a 240 scale
    assert hsv_of(255, 0, 0) == (0, 240, 240)
    assert hsv_of(255, 255, 0) == (40, 240, 240)
    assert hsv_of(0, 255, 0) == (80, 240, 240)
    assert hsv_of(0, 255, 255) == (120, 240, 240)
    assert hsv_of(0, 0, 255) == (160, 240, 240)
    assert hsv_of(255, 0, 255) == (200, 240, 240)
    # greys have no hue or saturation; V = floor(240 * x / 255)
    for x in (0, 1, 17, 128, 254, 255):
        assert hsv_of(x, x, x) == (0, 0, (240 * x) // 255)


def test_rgb_to_hsv_vs_colorsys():
    rng = np.random.default_rng(11)
    px = rng.integers(0, 256, (4000, 3))
    off = 0
    for r, g, b in px:
        h, s, v = hsv_of(int(r), int(g), int(b))
        ch, cs, cv = colorsys.rgb_to_hsv(r / 255, g / 255, b / 255)
        assert v == int(cv * 240)
        assert abs(s - int(cs * 240)) <= 1
        if s:
            # the reference wraps 360 degrees to 0 before scaling
            want = int((ch * 360 % 360) * (240.0 / 360.0))
            off += h != want
            assert abs(h - want) <= 1 or abs(h - want) == 239
    # the two evaluations round differently on a handful of exact ties only
    assert off < 40


def test_rgb_to_gray_unit_colours():
    rgb = np.array([[[1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, 0]]], np.uint8)
    rgb = np.repeat(rgb, 4, 0)
    gray, _ = orc_rgb_planes(rgb)
    assert np.array_equal(gray[0], [0.2989, 0.5870, 0.1140, 0.0])


def test_wide_rgb_scaling_identity():
    rng = np.random.default_rng(3)
    rgb8 = rng.integers(0, 256, (40, 36, 3), dtype=np.uint8)
    rgb8[0, 0] = (0, 0, 0)
    rgb8[0, 1] = (255, 255, 255)
    g8, h8 = orc_rgb_planes(rgb8)
    g16, h16 = orc_rgb_planes(rgb8.astype(np.uint16))
    assert np.array_equal(g8, g16) and np.array_equal(h8, h16)


def test_color_histogram_counts_classes():
    rgb = np.random.default_rng(8).integers(0, 256, (64, 48, 3), dtype=np.uint8)
    hist = orc_color_histogram(rgb)
    _, hsv = orc_rgb_planes(rgb)
    cls = orc_find_color_many(hsv)
    assert hist[0] == 0
    assert np.array_equal(hist, np.bincount(cls, minlength=20) / (64 * 48))
    # "Color": the class scaled to 0..255 and truncated; "Hue": the H byte
    col = orc_color_transform(T_COLOR, rgb).ravel()
    assert np.array_equal(col, np.floor(255.0 * cls / 19))
    assert np.array_equal(orc_color_transform(T_HUE, rgb), hsv[..., 0].astype(float))


def test_colour_plan_layout():
    lib = oracle()
    for s in range(4):
        assert lib.orc_plan_n_features(s) == SET_NFEAT[s]
    short, long_, short_c, long_c = (orc_names(s) for s in range(4))
    assert short_c[:len(short)] == short
    assert short_c[len(short)] == 'Color Histogram () [0]'
    it = iter(long_c)
    assert all(n in it for n in long_)  # long set's columns, in order
    groups = lambda names: [n.rsplit(' [', 1)[0] for n in names]
    extra = set(groups(long_c)) - set(groups(long_))
    assert extra == set(groups(short_c)) - set(groups(short))
    nodes = {g[g.index(' (') + 1:] for g in extra}
    assert nodes == {'()', '(Color Transform ())', '(Hue ())', '(Fourier (Hue ()))', '(Chebyshev (Hue ()))'}


@pytest.mark.parametrize('bad', ['Hue (Fourier ())', 'Color Histogram (Hue ())', 'Color (Color ())'])
def test_colour_group_validation(bad):
    """The colour transforms read the loaded image's colour plane, so they
    apply to the image itself; the Color Histogram reads it too."""
    from helpers import ROOT
    import os
    if not os.path.exists(os.path.join(ROOT, 'wnd-charm_amd', 'lib', 'libwndgpu.so')):
        pytest.skip('library not built')
    import wndcharm
    p = wndcharm.FeatureComputationPlan('bad')
    p.add('Pixel Intensity Statistics (%s)' % bad if not bad.startswith('Color Histogram') else bad)
    with pytest.raises(wndcharm.WndGpuError):
        p.finalize()
