"""GPU parity of sampled feature extraction (crop, Downsample, mean / stdev
normalisation, rotations, tiles) against the oracle, through
wndgpu_compute_sampled: every sample's feature vector at the usual bars
(integer-valued bit-exact, the rest within 1e-6 relative).  The pre-processed
planes feed min / max / median and every histogram, so an off-by-one pixel
or a rounding difference in the normalisation shows up as an exact-feature
mismatch."""
import numpy as np
import pytest

import wndcharm
from wndcharm import synthetic
from helpers import exact_mask_for_names, orc_features_sampled, orc_names
from test_gpu_parity import compare

pytestmark = pytest.mark.gpu


def run_case(img, set_, **opt):
    s = wndcharm.Sampling(**opt)
    plan = [wndcharm.StdFeatureComputationPlans.getFeatureSet, wndcharm.StdFeatureComputationPlans.getFeatureSetLong,
            wndcharm.StdFeatureComputationPlans.getFeatureSetColor][set_]()
    got = wndcharm.compute_features([img, img], plan, sampling=s)
    want = orc_features_sampled(img, set_, **opt)
    assert want is not None
    mask = exact_mask_for_names(orc_names(set_))
    ns = s.n_samples
    for k in range(ns):
        compare(got[k], want[k], mask, 'sample %d %s' % (k, opt))
        assert np.array_equal(got[ns + k], got[k])  # the second copy of the image


def test_crop_downsample_normalise_rotate_tile_u16():
    img = synthetic.image(3101, 10, (100, 110))
    print('sampled u16 crop+ds+norm+rot2+tiles2x2', flush=True)
    run_case(img, 0, roi=(12, 6, 84, 84), downsample=75, mean=1000.0, stdev=100.0, rotations=2, tiles=(2, 2))


def test_rotations_long_set():
    img = synthetic.image(3102, 8, (72, 72))
    print('sampled long set rotations 4', flush=True)
    run_case(img, 1, rotations=4)


def test_mean_only_u8_tiles():
    img = (synthetic.image(3103, 12, (96, 128)) // 16).astype(np.uint8)
    print('sampled u8 mean-only tiles 2x3', flush=True)
    run_case(img, 0, mean=90.0, tiles=(2, 3))


def test_colour_sampled():
    rng = np.random.default_rng(7)
    yy, xx = np.mgrid[0:90, 0:100]
    rgb = np.stack([(xx * 2.5) % 256, (yy * 2.8) % 256, ((xx + yy) * 1.3) % 256], -1)
    rgb = np.clip(rgb + rng.normal(0, 5, rgb.shape), 0, 255).astype(np.uint8)
    print('sampled colour crop+ds+rot2', flush=True)
    run_case(rgb, 2, roi=(4, 2, 80, 80), downsample=60, rotations=2)


def test_wide_rgb_crop_keeps_whole_image_range():
    rng = np.random.default_rng(8)
    rgb = rng.integers(2000, 30000, (64, 80, 3)).astype(np.uint16)
    rgb[0, 0] = (100, 100, 100)  # the range's minimum lies outside the crop
    print('sampled rgb16 crop', flush=True)
    run_case(rgb, 2, roi=(20, 16, 48, 40))


def test_refused_samples():
    plan = wndcharm.StdFeatureComputationPlans.getFeatureSet()
    img = synthetic.image(3104, 8, (60, 84))
    print('sampled refusals', flush=True)
    # a 90-degree rotation of a non-square image with tiles reaches past it
    _, st = wndcharm.compute_features([img], plan, return_status=True,
                                      sampling=wndcharm.Sampling(rotations=2, tiles=(3, 3)))
    assert list(st) == [5]
    _, st = wndcharm.compute_features([img.astype(np.float64)], plan, return_status=True,
                                      sampling=wndcharm.Sampling(mean=100.0))
    assert list(st) == [5]
    _, st = wndcharm.compute_features([img], plan, return_status=True,
                                      sampling=wndcharm.Sampling(roi=(50, 0, 40, 40)))
    assert list(st) == [5]
    _, st = wndcharm.compute_features([img], plan, return_status=True, sampling=wndcharm.Sampling(tiles=(6, 6)))
    assert list(st) == [1]  # 14 x 10 tiles are below the smallest side
