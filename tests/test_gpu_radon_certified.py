"""Radon features (radon.cpp:14-91 and RadonTextures, cmatrix.cpp:1463-1514
in the reference): the certified path sums each projection bin in any order
with a bound on its distance from the reference's sequential sum, settles
each angle's extreme bins and every 3-bin histogram slot from those bounds,
and recomputes in the reference's order the bins a decision leaves open.
Its 12 counts per plane must therefore equal those of the direct path (the
reference's order for every bin, WNDGPU_RADON_DIRECT=1) bit for bit, and
stay equal when WNDGPU_RADON_SLACK widens the bounds so that some (1e6) or
all (1e15) bins are recomputed.  The planes cover the nodes with negative
values (Chebyshev, Wavelet), a constant image (every interior bin of a
projection equal) and the sizes at which a block sums 4, 2 and 1 angles."""
import numpy as np
import pytest

import wndcharm
from wndcharm import synthetic
from helpers import ALG_NAMES, features_in_context, load_npz, orc_algorithm, orc_transform

pytestmark = pytest.mark.gpu
RADON = ALG_NAMES.index('Radon Coefficients')
NODES = ['()', '(Fourier ())', '(Chebyshev ())', '(Wavelet ())', '(Edge ())', '(Chebyshev (Fourier ()))',
         '(Wavelet (Edge ()))']


def _plan():
    plan = wndcharm.FeatureComputationPlan('radon')
    for node in NODES:
        plan.add('Radon Coefficients ' + node)
    plan.finalize()
    return plan


def _images(size):
    """'small': every plane's projection has at most 768 bins (4 angles per
    block); 'mid' adds a 700 x 700 image (991 bins, 2 angles); 'large' a
    1100 x 1100 one (1557 bins, 1 angle)"""
    if size == 'small':
        t1 = load_npz('img_t1_s01_c05_ij.npz')['pixels']
        return [t1, synthetic.image(3201, 6, (256, 256)), synthetic.image(3202, 24, (152, 445)),
                synthetic.image(3203, 12, (64, 64)), synthetic.image(3204, 48, (40, 97)),
                np.full((48, 48), 777, np.uint16)]
    side = {'mid': 700, 'large': 1100}[size]
    return [synthetic.image(3210, 12, (side, side)), synthetic.image(3211, 6, (96, 128))]


_direct = {}


def direct(size):
    if size not in _direct:
        print('radon: direct path', size, flush=True)
        _direct[size] = features_in_context({'WNDGPU_RADON_DIRECT': '1'}, _images(size), _plan())
    return _direct[size]


@pytest.mark.parametrize('size', ['small', 'mid', 'large'])
def test_certified_equals_direct(size):
    want = direct(size)
    print('radon: certified path', size, flush=True)
    got = features_in_context({}, _images(size), _plan())
    assert got.shape == want.shape == (len(_images(size)), 12 * len(NODES))
    assert np.array_equal(got, want, equal_nan=True)


@pytest.mark.parametrize('size,slack', [('small', 1e6), ('small', 1e15), ('mid', 1e15), ('large', 1e15)])
def test_widened_bounds_equal_direct(size, slack):
    """at 1e15 every bin is recomputed by the wave routine, on the larger
    planes over many 256-sub-column chunks and 512-term windows per bin"""
    want = direct(size)
    print('radon: certified path, slack', slack, size, flush=True)
    got = features_in_context({'WNDGPU_RADON_SLACK': repr(slack)}, _images(size), _plan())
    assert np.array_equal(got, want, equal_nan=True)


def test_direct_matches_oracle():
    """the direct path is the reference's arithmetic (its counts equal the
    oracle's)"""
    imgs = _images('small')
    got = direct('small')
    chains = {'()': [], '(Fourier ())': [0], '(Chebyshev ())': [1], '(Wavelet ())': [2], '(Edge ())': [3],
              '(Chebyshev (Fourier ()))': [0, 1], '(Wavelet (Edge ()))': [3, 2]}
    for k in (0, 3, 5):
        want = []
        for node in NODES:
            p = imgs[k].astype(np.float64)
            for t in chains[node]:
                p = orc_transform(t, p)
            want.append(orc_algorithm(RADON, p))
        assert np.array_equal(got[k], np.concatenate(want)), k
