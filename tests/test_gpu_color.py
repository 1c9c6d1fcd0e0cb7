"""GPU parity of the colour path against the oracle: the gray and colour
planes of RGB images (8- and 16-bit), the "Color" / "Hue" / "Color
Transform" transforms, the Color Histogram, and the short / long colour
feature sets end to end.  Plane values and the histogram are bit-exact;
feature bars as in test_gpu_parity (integer-valued bit-exact, the rest
within 1e-6 relative)."""
import numpy as np
import pytest

import wndcharm
from helpers import (T_COLOR, T_HUE, exact_mask_for_names, orc_color_histogram, orc_color_transform,
                     orc_features_rgb, orc_names, orc_rgb_planes)
from test_gpu_parity import compare

pytestmark = pytest.mark.gpu


def rgb_images():
    rng = np.random.default_rng(21)
    noise = rng.integers(0, 256, (96, 112, 3), dtype=np.uint8)
    # smooth colour gradients with texture: every hue and a spread of
    # saturations and values
    yy, xx = np.mgrid[0:80, 0:72]
    r = 127 + 120 * np.sin(xx / 7.0) * np.cos(yy / 11.0)
    g = 127 + 120 * np.sin((xx + yy) / 9.0)
    b = 127 + 120 * np.cos(xx / 5.0 - yy / 13.0)
    smooth = np.clip(np.stack([r, g, b], -1) + rng.normal(0, 6, (80, 72, 3)), 0, 255).astype(np.uint8)
    wide = (rng.integers(300, 40000, (64, 70, 3))).astype(np.uint16)
    return {'noise_u8': noise, 'smooth_u8': smooth, 'wide_u16': wide}


def cube(step):
    """every step-th value of each channel as one RGB image"""
    v = np.arange(0, 256, step)
    r, g, b = np.meshgrid(v, v, v, indexing='ij')
    n = len(v)
    return np.ascontiguousarray(np.stack([r, g, b], -1).reshape(n * n // 2 if n % 2 == 0 else n * n, -1, 3)
                                .astype(np.uint8))


@pytest.mark.parametrize('src', ['noise_u8', 'smooth_u8', 'wide_u16'])
def test_colour_planes_and_transforms(src):
    print('colour transforms', src, flush=True)
    rgb = rgb_images()[src]
    gray, _ = orc_rgb_planes(rgb)
    assert np.array_equal(wndcharm.ImageTransform('Color Transform').execute(rgb), gray)
    for name, kind in (('Color', T_COLOR), ('Hue', T_HUE)):
        got = wndcharm.ImageTransform(name).execute(rgb)
        assert np.array_equal(got, orc_color_transform(kind, rgb)), name


def test_colour_class_table_over_rgb_cube():
    """the device's class table against the oracle on every colour with
    channels in steps of 3 (86^3 pixels, ~4e5 distinct HSV triples)"""
    rgb = cube(3)
    print('colour cube', rgb.shape, flush=True)
    got = wndcharm.ImageTransform('Color').execute(rgb)
    assert np.array_equal(got, orc_color_transform(T_COLOR, rgb))


@pytest.mark.parametrize('src', ['noise_u8', 'smooth_u8', 'wide_u16'])
def test_color_histogram(src):
    print('colour histogram', src, flush=True)
    rgb = rgb_images()[src]
    plan = wndcharm.FeatureComputationPlan('hist')
    plan.add('Color Histogram ()')
    got = wndcharm.compute_features([rgb], plan)[0]
    assert np.array_equal(got, orc_color_histogram(rgb))


@pytest.mark.parametrize('set_', [2, 3])
def test_colour_sets(set_):
    imgs = rgb_images()
    names = orc_names(set_)
    plan = [wndcharm.StdFeatureComputationPlans.getFeatureSetColor,
            wndcharm.StdFeatureComputationPlans.getFeatureSetLongColor][set_ - 2]()
    assert plan.feature_names() == names
    print('colour set', set_, [im.shape for im in imgs.values()], flush=True)
    got = wndcharm.compute_features(list(imgs.values()), plan)
    mask = exact_mask_for_names(names)
    for k, (nm, im) in enumerate(imgs.items()):
        compare(got[k], orc_features_rgb(im, set_), mask, 'colour set %d %s' % (set_, nm))


def test_rgb_through_gray_sets():
    """an RGB image through the gray long set is its RGB2GRAY plane"""
    rgb = rgb_images()['smooth_u8']
    print('rgb through long set', flush=True)
    got = wndcharm.compute_features([rgb], wndcharm.StdFeatureComputationPlans.getFeatureSetLong())[0]
    compare(got, orc_features_rgb(rgb, 1), exact_mask_for_names(orc_names(1)), 'long set of rgb')


def test_gray_image_rejected_by_colour_plan():
    plan = wndcharm.StdFeatureComputationPlans.getFeatureSetColor()
    rgb = rgb_images()['smooth_u8']
    print('colour plan status', flush=True)
    out, st = wndcharm.compute_features([rgb, rgb[..., 0].copy(), rgb], plan, return_status=True)
    assert list(st) == [0, 4, 0]
    assert np.all(np.isnan(out[1]))
    assert np.array_equal(out[0], out[2])


def test_open_colour_tiff_through_executor(tmp_path):
    """ImageMatrix.OpenImage of an RGB TIFF keeps its samples; the executor
    runs the short colour set on them"""
    Image = pytest.importorskip('PIL.Image')
    rgb = rgb_images()['smooth_u8']
    p = str(tmp_path / 'c.tif')
    Image.fromarray(rgb).save(p, compression='tiff_lzw')
    print('open colour tiff', flush=True)
    im = wndcharm.ImageMatrix()
    assert im.OpenImage(p) == 1
    gray, _ = orc_rgb_planes(rgb)
    assert np.array_equal(im.as_ndarray(), gray)
    plan = wndcharm.StdFeatureComputationPlans.getFeatureSetColor()
    vec = wndcharm.DoubleVector(plan.n_features)
    wndcharm.FeatureComputationPlanExecutor(plan).run(im, vec, 0)
    compare(vec.as_ndarray(), orc_features_rgb(rgb, 2), exact_mask_for_names(orc_names(2)), 'colour tiff')
