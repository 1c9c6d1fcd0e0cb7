"""GPU parity on the benchmark configurations and the reference's own
end-to-end data, through the C ABI:

* the large-feature-set configuration (4 classes x 100 seeded 256x256
  images, periods 6/12/24/48, seed 2002): long-set features of two images
  per class against the oracle, and the Fisher-15% + WND5 step on the GPU
  features of all 400 images against the oracle classifier;
* planes above 300 x 300 pixels, whose Chebyshev-Fourier coefficients are
  taken after ImageMatrix::Downsample (src/cmatrix.cpp:1227-1240, 579-720):
  the reference's 301x300 golden image and two tiles of the uneven-tiles
  configuration (682x300, 514x448);
* the feature groups the standard plans compute on the root only (Edge,
  Otsu, Inverse-Otsu, Gabor) on transformed nodes of a custom plan, which
  FeatureNames::getGroupByName accepts (src/FeatureNames.cpp:309-396);
* the eight TIFFs of tests/wndchrm_tests/images through read_tiff, the long
  set and the .fit writer, against unbalanced.fit
  (tests/wndchrm_tests/test04-sigs.pl:16-31).

Bar: integer-valued features bit-exact, floating-point features within 1e-6
relative; classes bit-exact."""
import gzip
import os
import shutil
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import wndcharm
from wndcharm import synthetic
from helpers import (ALG_NAMES, EXACT, GOLDEN, exact_mask_for_names, load_npz, oracle, orc_algorithm, orc_features,
                     orc_names, orc_transform, ptr, wndcharm_compare)

pytestmark = pytest.mark.gpu
RTOL = 1e-6
POOL = 8


def compare_rows(got, want, mask, label):
    got = np.asarray(got, float)
    want = np.asarray(want, float)
    assert got.shape == want.shape, label
    both_nan = np.isnan(got) & np.isnan(want)
    ex_bad = mask & ~(got == want) & ~both_nan
    close = (got == want) | (np.abs(got - want) <= RTOL * np.abs(want)) | (np.abs(got - want) < 1e-12) | both_nan
    fp_bad = ~mask & ~close
    assert not ex_bad.any() and not fp_bad.any(), '%s: %d integer, %d floating-point mismatches, first %s' % (
        label, ex_bad.sum(), fp_bad.sum(), np.argwhere(ex_bad | fp_bad)[:6].tolist())


def oracle_rows(imgs, long_set=True):
    with ThreadPoolExecutor(POOL) as ex:
        return np.stack(list(ex.map(lambda im: orc_features(im.astype(np.float64), long_set), imgs)))


def test_long_set_large_config_images():
    """two images per class of the 4 x 100 256x256 configuration"""
    imgs, cls = synthetic.config_images('long_256')
    pick = [0, 57, 100, 163, 200, 251, 300, 399]
    sel = [imgs[i] for i in pick]
    print('long set, 256x256 configuration images', pick, flush=True)
    got = wndcharm.compute_features(sel, wndcharm.StdFeatureComputationPlans.getFeatureSetLong())
    want = oracle_rows(sel)
    mask = exact_mask_for_names(orc_names(True))
    for k, i in enumerate(pick):
        compare_rows(got[k], want[k], mask, 'image %d' % i)


def _oracle_classify(train, cls, n_classes, test, used):
    lib = oracle()
    tr = np.ascontiguousarray(train.copy())
    n, f = tr.shape
    mins, maxs = np.zeros(f), np.zeros(f)
    lib.orc_normalize_train(ptr(tr), n, f, ptr(mins), ptr(maxs))
    c = np.ascontiguousarray(cls, np.int32)
    w = np.zeros(f)
    lib.orc_fisher(ptr(tr), ptr(c), n, f, n_classes, used, ptr(w))
    probs, norms, preds = [], [], []
    for t in test:
        v = np.ascontiguousarray(t.copy())
        lib.orc_normalize_test(ptr(v), f, ptr(mins), ptr(maxs))
        p, nf = np.zeros(n_classes), np.zeros(1)
        preds.append(lib.orc_wnd5(ptr(tr), ptr(c), n, f, n_classes, ptr(w), ptr(v), ptr(p), ptr(nf)))
        probs.append(p)
        norms.append(nf[0])
    return tr, w, np.array(probs), np.array(norms), np.array(preds)


def test_large_config_fisher_wnd5():
    """GPU features of all 400 images; per class the first 75 train and the
    last 25 test; normalisation, Fisher weights keeping the top 15% and WND5
    on the GPU, against the oracle classifier on the same features"""
    imgs, cls = synthetic.config_images('long_256')
    print('long set, 400 images of 256x256, then Fisher 0.15 + WND5', flush=True)
    feats = wndcharm.compute_features(imgs, wndcharm.StdFeatureComputationPlans.getFeatureSetLong())
    # a few features are NaN by the reference's own arithmetic (e.g. ratios
    # over empty sets); normalisation maps them to 0 on both sides
    print('non-finite values: %d in columns %s' % ((~np.isfinite(feats)).sum(),
                                                  np.unique(np.where(~np.isfinite(feats))[1])[:10]), flush=True)
    pos = np.arange(400) % 100
    train, test = np.where(pos < 75)[0], np.where(pos >= 75)[0]
    tr, mins, maxs = wndcharm.normalize_train(feats[train])
    w = wndcharm.fisher_weights(tr, cls[train], 4, 0.15)
    te = np.stack([wndcharm.normalize_test(feats[t], mins, maxs) for t in test])
    prob, norm, pred = wndcharm.wnd5(tr, cls[train], 4, w, te)
    otr, ow, oprob, onorm, opred = _oracle_classify(feats[train], cls[train], 4, feats[test], 0.15)
    assert np.array_equal(tr, otr)
    assert np.array_equal(w, ow)
    assert (w > 0).sum() <= int(0.15 * 2919 + 0.5) + 1
    assert list(pred) == list(opred)
    np.testing.assert_allclose(prob, oprob, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(norm, onorm, rtol=1e-9)
    acc = np.mean(pred == cls[test])
    print('accuracy on the held-out 100 images: %.3f' % acc, flush=True)
    assert acc > 0.9


@pytest.mark.parametrize('case', ['golden_301x300', 'tile0_682x300', 'tile3_514x448'])
def test_long_set_above_300x300(case):
    """long set on planes whose Chebyshev-Fourier input is downsampled"""
    if case == 'golden_301x300':
        img = load_npz('img_010067_301x300.npz')['pixels']
    else:
        tiles, _ = synthetic.config_images('tiles_mixed', limit=4)
        img = tiles[int(case[4])]
    assert img.shape[0] * img.shape[1] > 300 * 300
    print('long set above 300x300:', case, img.shape, flush=True)
    got = wndcharm.compute_features([img], wndcharm.StdFeatureComputationPlans.getFeatureSetLong())[0]
    want = orc_features(img.astype(np.float64), True)
    names = orc_names(True)
    compare_rows(got, want, exact_mask_for_names(names), case)
    # the downsampled group itself on the root and on the transformed nodes
    cf = [i for i, n in enumerate(names) if n.startswith('Chebyshev-Fourier Coefficients')]
    assert np.array_equal(got[cf], want[cf])


GROUP_A = [10, 11, 12, 13]  # Edge Features, Otsu / Inverse-Otsu Object Features, Gabor Textures
NODES = {'Fourier ()': 0, 'Wavelet ()': 2, 'Chebyshev ()': 1, 'Edge ()': 3}


@pytest.mark.parametrize('src', ['t1_113', 'synth_128'])
def test_group_a_on_transformed_nodes_custom_plan(src):
    """Edge, Otsu, Inverse-Otsu and Gabor groups on the F, W, C and E nodes,
    16 groups in one custom plan, against the oracle algorithm on the
    oracle's transformed plane"""
    if src == 't1_113':
        img = load_npz('img_t1_s01_c05_ij.npz')['pixels'].astype(np.float64)
    else:
        img = synthetic.image(1001, 8, (128, 128)).astype(np.float64)
    plan = wndcharm.FeatureComputationPlan('group A on transformed nodes')
    groups = ['%s (%s)' % (ALG_NAMES[a], node) for node in NODES for a in GROUP_A]
    for g in groups:
        plan.add(g)
    plan.finalize()
    print('custom plan of', len(groups), 'groups on', src, flush=True)
    got = wndcharm.compute_features([img], plan)[0]
    for node, tid in NODES.items():
        plane = orc_transform(tid, img)
        for a in GROUP_A:
            g = '%s (%s)' % (ALG_NAMES[a], node)
            off = plan.getFGoffset(g)
            want = orc_algorithm(a, plane)
            mask = np.zeros(len(want), bool)
            mask[EXACT[a]] = True
            compare_rows(got[off:off + len(want)], want, mask, g)


TIFFS = ['binuc1', 'binuc2', 'binuc3', 'binuc4', 'binuc5', 'norm1', 'norm2', 'norm3']


def f14_column(name):
    return name.startswith('Haralick Textures') and (name.endswith('[16]') or name.endswith('[17]'))


def test_unbalanced_fit_end_to_end():
    """`wndchrm train -l images` of test04-sigs.pl on the engine: the eight
    TIFFs read by read_tiff, long-set features in one GPU batch, written by
    the .fit writer, then compared with unbalanced.fit.  Headers, names,
    class labels, classes and paths must be identical and every value equal
    under the reference's comparator (wndcharm/utils.py:276-366), except the
    two f14 columns of each Haralick group: unbalanced.fit predates the
    reference's current f14 (CVIPtexture.cpp:993-1023), which the GPU
    follows bit for bit against the reference source itself."""
    tmp = tempfile.mkdtemp()
    try:
        imgs = []
        for t in TIFFS:
            path = os.path.join(tmp, t + '.tiff')
            with gzip.open(os.path.join(GOLDEN, 'raw', t + '.tiff.gz'), 'rb') as fi, open(path, 'wb') as fo:
                fo.write(fi.read())
            imgs.append(wndcharm.read_tiff(path))
        ref = load_npz('unbalanced_fit.npz')
        for k, im in enumerate(imgs):
            assert np.array_equal(im, ref['pixels'][k]), TIFFS[k]
        print('unbalanced.fit end to end: 8 TIFFs', imgs[0].shape, flush=True)
        plan = wndcharm.StdFeatureComputationPlans.getFeatureSetLong()
        got = wndcharm.compute_features(imgs, plan)
        golden_path = os.path.join(tmp, 'unbalanced.fit')
        with gzip.open(os.path.join(GOLDEN, 'raw', 'unbalanced.fit.gz'), 'rb') as fi, open(golden_path, 'wb') as fo:
            fo.write(fi.read())
        gold = wndcharm.read_fit(golden_path)
        # the .fit lists features by name; map the plan's columns onto them
        col = {n: i for i, n in enumerate(plan.feature_names())}
        order = [col[n] for n in gold.names]
        out = os.path.join(tmp, 'test-images-sigs-l.fit')
        wndcharm.write_fit(out, got[:, order], gold.classes, gold.names, gold.class_labels, gold.paths,
                           version=gold.version, vec_type=gold.feature_vec_type)
        mine_lines = open(out).read().splitlines()
        gold_lines = open(golden_path).read().splitlines()
        assert len(mine_lines) == len(gold_lines)
        nf = len(gold.names)
        head = 3 + nf + 1 + gold.n_classes
        assert mine_lines[:head] == gold_lines[:head]
        bad, same_text, compared = [], 0, 0
        for s in range(8):
            a = mine_lines[head + 2 * s].split()
            b = gold_lines[head + 2 * s].split()
            assert len(a) == len(b) == nf + 1
            assert a[nf] == b[nf]  # class
            assert mine_lines[head + 2 * s + 1] == gold_lines[head + 2 * s + 1]  # path
            for j in range(nf):
                if f14_column(gold.names[j]):
                    continue
                compared += 1
                same_text += a[j] == b[j]
                if not wndcharm_compare(float(a[j]), float(b[j])):
                    bad.append((TIFFS[s], gold.names[j], a[j], b[j]))
        print('values compared: %d, textually identical: %d' % (compared, same_text), flush=True)
        assert not bad, '%d values differ, first %s' % (len(bad), bad[:6])
        # and the file reads back to the computed values at %.8g
        back = wndcharm.read_fit(out)
        np.testing.assert_allclose(back.values, got[:, order], rtol=1e-7, atol=0)
    finally:
        shutil.rmtree(tmp)


@pytest.mark.parametrize('shape', [(16, 16), (17, 23), (31, 16), (16, 40)])
def test_long_set_smallest_planes(shape):
    """the long set on planes at the smallest accepted sides (16): tiny
    Haralick tone counts and Chebyshev-Fourier orders (N = (min side - 1) / 2),
    short FFT lines with prime lengths, wavelet nodes below 16 pixels,
    against the oracle (integer-valued columns exact, the rest to 1e-6)"""
    img = synthetic.image(7000 + shape[0] * 100 + shape[1], 5, shape)
    print('long set on', shape, flush=True)
    got = wndcharm.compute_features([img], wndcharm.StdFeatureComputationPlans.getFeatureSetLong())[0]
    want = orc_features(img.astype(np.float64), True)
    compare_rows(got, want, exact_mask_for_names(orc_names(True)), str(shape))
