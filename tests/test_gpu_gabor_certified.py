"""Gabor textures of the plain form (GaborTextureFilters2D, gabor.cpp:201-231
in the reference): the energies are convolved with FP64 fused multiply-adds,
each with an a-priori bound on its distance from the reference's separate
multiply-and-add value, and every count is decided from those bounds, the
energies a decision leaves open being recomputed in the reference's order.
The ratios must therefore equal those of the direct path (the reference's
rounding throughout, WNDGPU_GABOR_DIRECT=1) bit for bit, and stay equal when
WNDGPU_GABOR_SLACK widens the bounds so that many energies are recomputed,
by the per-pixel list repair or, past 63 per filter, tile by tile."""
import numpy as np
import pytest

import wndcharm
from wndcharm import synthetic
from helpers import ALG_NAMES, features_in_context, load_npz, orc_algorithm, orc_transform

pytestmark = pytest.mark.gpu
GABOR = ALG_NAMES.index('Gabor Textures')
NODES = ['()', '(Fourier ())', '(Chebyshev ())', '(Wavelet ())', '(Edge ())']


def _plan():
    plan = wndcharm.FeatureComputationPlan('gabor')
    for node in NODES:
        plan.add('Gabor Textures ' + node)
    plan.finalize()
    return plan


def _images():
    t1 = load_npz('img_t1_s01_c05_ij.npz')['pixels']
    imgs = [t1, synthetic.image(3101, 6, (256, 256)), synthetic.image(3102, 24, (152, 445)),
            synthetic.image(3103, 12, (64, 64)), synthetic.image(3104, 48, (40, 97))]
    imgs.append(np.full((48, 48), 777, np.uint16))  # constant: every energy equal
    return imgs


def _features(env):
    """features of _images() under _plan() in a fresh context created with env"""
    return features_in_context(env, _images(), _plan())


@pytest.fixture(scope='module')
def direct():
    print('gabor: direct path', flush=True)
    return _features({'WNDGPU_GABOR_DIRECT': '1'})


def test_certified_equals_direct(direct):
    print('gabor: certified path', flush=True)
    got = _features({})
    assert got.shape == direct.shape == (6, 7 * len(NODES))
    assert np.array_equal(got, direct, equal_nan=True)


@pytest.mark.parametrize('slack', [1e3, 1e12])
def test_widened_bounds_equal_direct(direct, slack):
    """1e3: a few energies per filter open (list repair); 1e12: most of them
    (tile repair)"""
    print('gabor: certified path, slack', slack, flush=True)
    got = _features({'WNDGPU_GABOR_SLACK': repr(slack)})
    assert np.array_equal(got, direct, equal_nan=True)


def test_direct_matches_oracle(direct):
    """the direct path is the reference's arithmetic (its counts equal the
    oracle's on every node)"""
    imgs = _images()
    chains = {'()': [], '(Fourier ())': [0], '(Chebyshev ())': [1], '(Wavelet ())': [2], '(Edge ())': [3]}
    for k in (0, 3):
        want = []
        for node in NODES:
            p = imgs[k].astype(np.float64)
            for t in chains[node]:
                p = orc_transform(t, p)
            want.append(orc_algorithm(GABOR, p))
        assert np.array_equal(direct[k], np.concatenate(want)), k
