"""The reduced-precision forms of the basis contractions (FP32, and BF16x3
split products on the bf16 matrix cores with FP32 accumulation) against the
FP64 oracle on the short-set 128x128 images: Zernike (an FP feature) within
1e-5 relative; Chebyshev, Chebyshev-Fourier and Gabor features are counts
of FP values in histogram bins or above an Otsu threshold, so they may move
only where a value lies on a bin edge -- at most a few values per family,
each by at most two counts."""
import numpy as np
import pytest

import wndcharm
from wndcharm import synthetic
from helpers import ALG_NAMES, orc_algorithm

pytestmark = pytest.mark.gpu
FAMILIES = ['Chebyshev Coefficients', 'Zernike Coefficients', 'Chebyshev-Fourier Coefficients', 'Gabor Textures']


@pytest.mark.parametrize('precision', [wndcharm.FP32, wndcharm.BF16X3])
def test_reduced_precision_contractions(precision):
    imgs, _ = synthetic.config_images('short_128')
    imgs = imgs[::2]
    plan = wndcharm.FeatureComputationPlan('contractions')
    for fam in FAMILIES:
        plan.add(fam + ' ()')
    plan.finalize()
    plan.set_precision(precision)
    print('reduced precision', precision, len(imgs), flush=True)
    got = wndcharm.compute_features(imgs, plan)
    off = 0
    for fam in FAMILIES:
        alg = ALG_NAMES.index(fam)
        want = np.stack([orc_algorithm(alg, im.astype(np.float64)) for im in imgs])
        n = want.shape[1]
        g = got[:, off:off + n]
        off += n
        if fam == 'Zernike Coefficients':
            assert np.allclose(g, want, rtol=1e-5, atol=0), np.abs(g / want - 1).max()
            continue
        changed = g != want
        assert changed.sum() <= 4, (fam, changed.sum())
        if fam == 'Gabor Textures':
            # ratios of counts: the base count of the lowest filter is ~1e3 pixels
            assert np.all(np.abs(g - want) <= 3e-3), (fam, np.abs(g - want).max())
        else:
            assert np.all(np.abs(g - want) <= 2), (fam, np.abs(g - want).max())
            assert np.array_equal(g.sum(1), want.sum(1))  # every coefficient is still counted once


def test_standard_plans_stay_plain():
    with pytest.raises(wndcharm.WndGpuError):
        wndcharm.StdFeatureComputationPlans.getFeatureSetLong().set_precision(wndcharm.BF16X3)
