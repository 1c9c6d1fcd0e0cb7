"""GPU parity: the HIP engine (through the C ABI, lib/libwndgpu.so) against
the CPU oracle on the same inputs.  Integer-valued features must be
bit-exact; floating-point features within 1e-6 relative (the north-star
bar) -- in practice they agree to ~1e-12."""
import numpy as np
import pytest

import wndcharm
from wndcharm import synthetic
from helpers import (ALG_NAMES, EXACT, exact_mask_for_names, load_npz, orc_algorithm, orc_features, orc_names,
                     orc_transform)

pytestmark = pytest.mark.gpu
RTOL = 1e-6


def compare(got, want, exact_mask, label):
    got = np.asarray(got, float)
    want = np.asarray(want, float)
    assert got.shape == want.shape, label
    bad = []
    for k in range(len(want)):
        g, w = got[k], want[k]
        if np.isnan(w) and np.isnan(g):
            continue
        if exact_mask[k]:
            if g != w:
                bad.append((k, g, w, 'exact'))
        else:
            if not (g == w or abs(g - w) <= RTOL * max(abs(w), 1e-300) or abs(g - w) < 1e-12):
                bad.append((k, g, w, 'rel %.2e' % (abs(g - w) / max(abs(w), 1e-300))))
    assert not bad, '%s: %d mismatches, first %s' % (label, len(bad), bad[:6])


def root_images():
    t1 = load_npz('img_t1_s01_c05_ij.npz')['pixels'].astype(np.float64)
    s = synthetic.image(1001, 8, (128, 128)).astype(np.float64)
    return {'t1_113': t1, 'synth_128': s}


@pytest.mark.parametrize('tname,tid', [('Edge', 3), ('Wavelet', 2), ('Chebyshev', 1), ('Fourier', 0)])
@pytest.mark.parametrize('shape', [(128, 128), (113, 113), (96, 80), (70, 100)])
def test_transform(tname, tid, shape):
    print('transform', tname, shape, flush=True)
    img = synthetic.image(7, 12, shape).astype(np.float64)
    got = wndcharm.ImageTransform(tname).execute(img)
    want = orc_transform(tid, img)
    assert got.shape == want.shape
    # all four transforms evaluate the oracle's operations in the oracle's order
    assert np.array_equal(got, want), np.abs(got - want).max()


@pytest.mark.parametrize('node', ['root', 'F', 'W', 'C', 'E'])
@pytest.mark.parametrize('src', ['t1_113', 'synth_128'])
@pytest.mark.parametrize('alg', list(range(15)))
def test_algorithm(alg, src, node):
    print('algorithm', ALG_NAMES[alg], src, node, flush=True)
    img = root_images()[src]
    if node != 'root':
        img = orc_transform({'F': 0, 'C': 1, 'W': 2, 'E': 3}[node], img)
    # Edge, Otsu, Inverse-Otsu and Gabor run on the root only in the standard
    # plans; custom plans may put them on any node
    got = wndcharm.FeatureAlgorithm(ALG_NAMES[alg]).execute(img)
    want = orc_algorithm(alg, img)
    mask = np.zeros(len(want), bool)
    mask[EXACT[alg]] = True
    compare(got, want, mask, '%s on %s %s' % (ALG_NAMES[alg], src, node))


@pytest.mark.parametrize('node', ['root', 'F', 'W', 'C', 'E'])
@pytest.mark.parametrize('src', ['t1_113', 'synth_128'])
def test_haralick_f14_bit_exact(src, node):
    """Haralick f14 (mean and range over the angles of sqrt(eigenvalue 2 of
    Q)) equals the oracle bit for bit: Q, the Hessenberg reduction and the
    Francis QR run in the reference's Eigen orders.  (f1..f13 use parallel
    sums and are held to 1e-6 by test_algorithm.)"""
    print('haralick exact', src, node, flush=True)
    img = root_images()[src]
    if node != 'root':
        img = orc_transform({'F': 0, 'C': 1, 'W': 2, 'E': 3}[node], img)
    got = np.asarray(wndcharm.FeatureAlgorithm('Haralick Textures').execute(img))
    want = orc_algorithm(ALG_NAMES.index('Haralick Textures'), img)
    f14 = [16, 17]
    assert np.array_equal(got[f14], want[f14]), (got[f14], want[f14])


def test_long_set_batch_uneven_sizes():
    imgs = [load_npz('img_t1_s01_c05_ij.npz')['pixels'], synthetic.image(2002, 6, (128, 128)),
            synthetic.image(2003, 24, (96, 140))]
    print('long set batch', [i.shape for i in imgs], flush=True)
    got = wndcharm.compute_features(imgs, wndcharm.StdFeatureComputationPlans.getFeatureSetLong())
    names = orc_names(True)
    assert wndcharm.StdFeatureComputationPlans.getFeatureSetLong().feature_names() == names
    mask = exact_mask_for_names(names)
    for k, im in enumerate(imgs):
        want = orc_features(im.astype(np.float64), True)
        compare(got[k], want, mask, 'long set image %d %s' % (k, im.shape))


def test_long_set_uneven_tiles():
    """tiles of the uneven-tiles configuration (sides uniform in [64, 1024]),
    batched with a square image"""
    tiles, _ = synthetic.config_images('tiles_mixed', limit=7)
    imgs = [tiles[6], tiles[2], synthetic.image(2004, 12, (64, 64))]
    print('long set uneven tiles', [i.shape for i in imgs], flush=True)
    got = wndcharm.compute_features(imgs, wndcharm.StdFeatureComputationPlans.getFeatureSetLong())
    mask = exact_mask_for_names(orc_names(True))
    for k, im in enumerate(imgs):
        compare(got[k], orc_features(im.astype(np.float64), True), mask, 'tile %d %s' % (k, im.shape))


def test_short_set_golden_tile():
    d = load_npz('img_lymphoma_tile.npz')
    print('short set lymphoma tile', d['pixels'].shape, flush=True)
    got = wndcharm.compute_features([d['pixels']], wndcharm.StdFeatureComputationPlans.getFeatureSet())[0]
    want = orc_features(d['pixels'].astype(np.float64), False)
    compare(got, want, exact_mask_for_names(orc_names(False)), 'short set tile')


def test_executor_api_matches_batch():
    img = synthetic.image(1005, 20, (128, 128))
    plan = wndcharm.StdFeatureComputationPlans.getFeatureSet()
    vec = wndcharm.DoubleVector(plan.n_features * 2)
    im = wndcharm.ImageMatrix(img.astype(np.float64))
    wndcharm.FeatureComputationPlanExecutor(plan).run(im, vec, 1)
    row = wndcharm.compute_features([img], plan)[0]
    assert np.array_equal(vec.as_ndarray()[plan.n_features:], row)
    assert np.all(vec.as_ndarray()[:plan.n_features] == 0)


def test_custom_plan():
    plan = wndcharm.FeatureComputationPlan('custom')
    plan.add('Zernike Coefficients (Wavelet (Edge ()))')
    plan.add('Haralick Textures (Chebyshev (Fourier ()))')
    plan.finalize()
    assert plan.n_features == 100
    assert plan.getFeatureNameByIndex(72) == 'Haralick Textures (Chebyshev (Fourier ())) [0]'
    img = synthetic.image(1009, 8, (128, 128)).astype(np.float64)
    got = wndcharm.compute_features([img], plan)[0]
    full = orc_features(img, True)
    names = orc_names(True)
    want = np.array([full[names.index(plan.getFeatureNameByIndex(i))] for i in range(100)])
    mask = exact_mask_for_names([plan.getFeatureNameByIndex(i) for i in range(100)])
    compare(got, want, mask, 'custom plan')


def test_bad_image_status():
    plan = wndcharm.StdFeatureComputationPlans.getFeatureSet()
    good = synthetic.image(1001, 8, (64, 64))
    out, st = wndcharm.compute_features([good, np.zeros((8, 8), np.uint16), good], plan, return_status=True)
    assert list(st) == [0, 1, 0]
    assert np.all(np.isnan(out[1]))
    assert np.array_equal(out[0], out[2])


def test_haralick_large_batch_equals_small_batches():
    """the Schur kernel runs one wavefront per matrix above 1024 matrices
    (256 images) and one 256-thread block per matrix below; both apply the
    same operations to every entry, so the 28 Haralick values of 260 images
    computed in one call equal those computed in two calls of 130 (the
    small-batch form is the one the oracle tests above run)"""
    imgs = [synthetic.image(5000 + k, 4 + k % 9, (64, 64 + 2 * (k % 5))) for k in range(260)]
    plan = wndcharm.FeatureComputationPlan('haralick')
    for g in ('Haralick Textures ()', 'Haralick Textures (Fourier ())'):
        plan.add(g)
    plan.finalize()
    print('haralick 260 vs 2 x 130', flush=True)
    big = wndcharm.compute_features(imgs, plan)
    small = np.concatenate([wndcharm.compute_features(imgs[:130], plan), wndcharm.compute_features(imgs[130:], plan)])
    assert np.array_equal(big, small, equal_nan=True)


HARALICK_NODES = ['()', '(Chebyshev ())', '(Chebyshev (Fourier ()))', '(Chebyshev (Wavelet ()))', '(Edge ())',
                  '(Fourier ())', '(Fourier (Chebyshev ()))', '(Fourier (Edge ()))', '(Fourier (Wavelet ()))',
                  '(Wavelet ())', '(Wavelet (Edge ()))', '(Wavelet (Fourier ()))']


def test_haralick_dispatch_order_beyond_one_chunk():
    """the eigen kernels take their matrices largest first through an order
    that a one-block counting sort builds from the entries' tone counts, its
    keys staged in LDS 16,384 entries at a time.  1,400 images x the long
    set's 12 Haralick nodes are 16,800 entries, two chunks: every entry must
    still be computed exactly once, so one call equals two calls of 700
    (8,400 entries, one chunk) bit for bit"""
    imgs = [synthetic.image(6100 + k, 3 + k % 11, (24 + k % 9, 24 + (k // 9) % 7)) for k in range(1400)]
    plan = wndcharm.FeatureComputationPlan('haralick, 12 nodes')
    for node in HARALICK_NODES:
        plan.add('Haralick Textures ' + node)
    plan.finalize()
    print('haralick 1400 x 12 nodes vs 2 x 700', flush=True)
    big = wndcharm.compute_features(imgs, plan)
    small = np.concatenate([wndcharm.compute_features(imgs[:700], plan), wndcharm.compute_features(imgs[700:], plan)])
    assert big.shape == (1400, 12 * 28) and np.isfinite(big).any()
    assert np.array_equal(big, small, equal_nan=True)
