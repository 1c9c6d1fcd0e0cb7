"""The reduced-precision forms (FP32, and BF16 split products on the bf16
matrix cores) against the FP64 oracle: Zernike (an FP feature) within 1e-5
relative; the integer-valued features -- Chebyshev and Chebyshev-Fourier
histogram counts, Gabor ratios of pixel counts -- bit-exact.  The Gabor
energies carry a-priori error bounds and every undecided count is settled on
an FP64 recomputation; the two histogram families keep their contractions in
FP64."""
import numpy as np
import pytest

import wndcharm
from wndcharm import synthetic
from helpers import ALG_NAMES, exact_mask_for_names, orc_algorithm, orc_transform

pytestmark = pytest.mark.gpu
FAMILIES = ['Chebyshev Coefficients', 'Zernike Coefficients', 'Chebyshev-Fourier Coefficients', 'Gabor Textures']


def _plan(groups, precision):
    plan = wndcharm.FeatureComputationPlan('contractions')
    for g in groups:
        plan.add(g)
    plan.finalize()
    plan.set_precision(precision)
    return plan


@pytest.mark.parametrize('precision', [wndcharm.FP32, wndcharm.BF16X3])
def test_reduced_precision_contractions(precision):
    imgs, _ = synthetic.config_images('short_128')
    plan = _plan([fam + ' ()' for fam in FAMILIES], precision)
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
        else:
            assert np.array_equal(g, want), (fam, int((g != want).sum()))


@pytest.mark.parametrize('precision', [wndcharm.FP32, wndcharm.BF16X3])
def test_reduced_precision_gabor_on_transformed_plane(precision):
    """Gabor on the Fourier node: pixels that are not integers (the BF16
    split of the input is then inexact and enters the bound)"""
    img = synthetic.image(1003, 8, (96, 112)).astype(np.float64)
    plan = _plan(['Gabor Textures (Fourier ())', 'Gabor Textures ()'], precision)
    print('gabor on the Fourier plane', precision, flush=True)
    got = wndcharm.compute_features([img], plan)[0]
    want_f = orc_algorithm(ALG_NAMES.index('Gabor Textures'), orc_transform(0, img))
    want_r = orc_algorithm(ALG_NAMES.index('Gabor Textures'), img)
    off_f = plan.getFGoffset('Gabor Textures (Fourier ())')
    off_r = plan.getFGoffset('Gabor Textures ()')
    assert np.array_equal(got[off_f:off_f + 7], want_f)
    assert np.array_equal(got[off_r:off_r + 7], want_r)


def test_long_set_bf16_integer_columns_exact():
    """The whole long set with the contractions in BF16 on 256x256
    configuration images and an uneven tile: every integer-valued column
    equals the FP64 form's, the FP columns are within 1e-5"""
    imgs, _ = synthetic.config_images('long_256')
    tiles, _ = synthetic.config_images('tiles_mixed', limit=7)
    sel = [imgs[3], imgs[150], imgs[377], tiles[6]]
    names = wndcharm.StdFeatureComputationPlans.getFeatureSetLong().feature_names()
    groups = []
    for nm in names:
        g = nm.rsplit(' [', 1)[0]
        if g not in groups:
            groups.append(g)
    plan = _plan(groups, wndcharm.BF16X3)
    assert plan.feature_names() == names
    print('long set in BF16', [s.shape for s in sel], flush=True)
    got = wndcharm.compute_features(sel, plan)
    ref = wndcharm.compute_features(sel, wndcharm.StdFeatureComputationPlans.getFeatureSetLong())
    exact = exact_mask_for_names(names)
    assert np.array_equal(got[:, exact], ref[:, exact], equal_nan=True), int((got[:, exact] != ref[:, exact]).sum())
    close = (got == ref) | (np.abs(got - ref) <= 1e-5 * np.abs(ref)) | (np.isnan(got) & np.isnan(ref))
    assert close.all(), np.argwhere(~close)[:5]


def test_standard_plans_stay_plain():
    with pytest.raises(wndcharm.WndGpuError):
        wndcharm.StdFeatureComputationPlans.getFeatureSetLong().set_precision(wndcharm.BF16X3)
