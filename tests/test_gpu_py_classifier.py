"""The wndcharm Python package's classifier on the GPU (wndgpu_py_normalize,
wndgpu_py_fisher, wndgpu_py_wnd5 behind wndcharm.utils / FeatureSpace /
FeatureVector / FeatureWeights / SingleSamplePrediction):

* the reference's own Python tests, written the way they are written there:
  test_SingleSamplePrediction.py:57-91 (WND5 marginal probabilities at
  2189 / 438 / 146 features), test_FeatureWeights.py:51-79 (Fisher weights
  against test_fit-l.weights, slicing), test_FeatureSpace.py:724-745
  (Normalize against test_fit-l-normalized.fit, z-scores against scipy);
* GPU against the oracle's restatement on the same inputs: bit-exact for
  normalisation and Fisher weights, probabilities within 1e-12 relative and
  identical predicted classes for WND5 (the GPU's pow may differ from the
  host's in the last bit);
* FeatureVector.GenerateFeatures of the reference's t1 TIFF against its
  .sig golden (reference comparator, f14 columns excluded as in the oracle's
  golden tests)."""
import ctypes
import gzip
import os
import shutil

import numpy as np
import pytest

import wndcharm
from wndcharm.FeatureSpace import FeatureSpace
from wndcharm.FeatureVector import FeatureVector
from wndcharm.FeatureWeights import FisherFeatureWeights
from wndcharm.SingleSamplePrediction import SingleSampleClassification
from wndcharm.FeatureSpacePrediction import FeatureSpaceClassification
from wndcharm.utils import ReplaceNonReal, normalize_by_columns
from helpers import GOLDEN, wndcharm_compare
import test_oracle_py_classifier as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def test_dir(tmp_path_factory):
    d = tmp_path_factory.mktemp('wndchrm_tests')
    for name in ('test-l.fit', 'test_fit-l.weights', 'test_fit-l-normalized.fit', 't1_s01_c05_ij-l_precalculated.sig',
                 't1_s01_c05_ij.tif'):
        with gzip.open(os.path.join(GOLDEN, 'raw', name + '.gz'), 'rb') as fi, open(str(d / name), 'wb') as fo:
            shutil.copyfileobj(fi, fo)
    return str(d)


def test_WND5_all_features(test_dir):
    """tests/pywndcharm_tests/test_SingleSamplePrediction.py:57-91"""
    epsilon = 0.00001
    correct_marg_probs = {2189: [0.083, 0.917], 438: [0.076, 0.924], 146: [0.066, 0.934]}
    feature_set = FeatureSpace.NewFromFitFile(os.path.join(test_dir, 'test-l.fit'))
    fs1 = feature_set.feature_names
    feature_set.Normalize()
    assert fs1 == feature_set.feature_names
    test_sample = FeatureVector(source_filepath=os.path.join(test_dir, 't1_s01_c05_ij.tif'), long=True)
    test_sample.LoadSigFile(os.path.join(test_dir, 't1_s01_c05_ij-l_precalculated.sig'))
    assert feature_set.feature_names == test_sample.feature_names
    test_sample.Normalize(feature_set)
    all_weights = FisherFeatureWeights.NewFromFile(os.path.join(test_dir, 'test_fit-l.weights'))
    for num_feats, want in correct_marg_probs.items():
        print('WND5 at %d features' % num_feats, flush=True)
        weights = all_weights.Threshold(num_feats)
        feat_set = feature_set.FeatureReduce(weights)
        sample = test_sample.FeatureReduce(weights)
        result = SingleSampleClassification.NewWND5(feat_set, weights, sample)
        got = [round(v, 3) for v in result.marginal_probabilities]
        for t, r in zip(want, got):
            assert abs(t - r) <= epsilon, (num_feats, got)
        assert result.predicted_label == '4cell'


def test_NewFromFeatureSet(test_dir):
    """tests/pywndcharm_tests/test_FeatureWeights.py:51-79"""
    fs = FeatureSpace.NewFromFitFile(os.path.join(test_dir, 'test-l.fit')).Normalize(inplace=True)
    fw = FisherFeatureWeights.NewFromFeatureSpace(fs)
    target_fw = FisherFeatureWeights.NewFromFile(os.path.join(test_dir, 'test_fit-l.weights'))
    for t, r in zip(target_fw.values, fw.values):
        assert abs(t - r) <= 0.00001
    sliced = fw[:10]
    assert len(sliced) == 10 and len(sliced.feature_names) == 10 and len(sliced.values) == 10
    for i in range(10):
        assert sliced.feature_names[i] == fw.feature_names[i] and sliced.values[i] == fw.values[i]
    sliced = fw[50:100:2]
    for i, j in zip(range(len(sliced)), range(50, 100, 2)):
        assert sliced.feature_names[i] == fw.feature_names[j] and sliced.values[i] == fw.values[j]


def test_Normalize(test_dir):
    """tests/pywndcharm_tests/test_FeatureSpace.py:724-745 (z-scores on the
    columns whose values are not all equal; see the oracle's z-score test)"""
    raw_fs = FeatureSpace.NewFromFitFile(os.path.join(test_dir, 'test-l.fit'))
    result_fs = raw_fs.Normalize(inplace=False)
    target_fs = FeatureSpace.NewFromFitFile(os.path.join(test_dir, 'test_fit-l-normalized.fit'))
    np.testing.assert_allclose(result_fs.data_matrix, target_fs.data_matrix, rtol=1e-05)
    from scipy.stats.mstats import zscore
    raw = raw_fs.data_matrix.copy()
    ReplaceNonReal(raw)
    with np.errstate(all='ignore'):
        target = np.asarray(zscore(raw))
    target[np.isnan(target)] = 0
    result_fs = raw_fs.Normalize(inplace=False, zscore=True)
    const = np.ptp(raw, axis=0) == 0
    np.testing.assert_allclose(result_fs.data_matrix[:, ~const], target[:, ~const], rtol=1e-7, atol=1e-12)


# ---------------------------------------------------------------- GPU = oracle

def _edge_matrix():
    rng = np.random.default_rng(21)
    M = rng.normal(0, 3, (37, 11))
    M[3, 0] = np.inf
    M[5, 1] = -np.inf
    M[7, 2] = np.nan
    M[:, 4] = 2.5                 # constant column
    M[:, 5] = np.nan              # no finite entry
    M[10, 6] = np.inf
    M[11, 6] = np.nan
    return M


@pytest.mark.parametrize('case', ['minmax', 'zscore', 'given_ranges', 'given_ranges_no_check', 'replace_only',
                                  'replace_given', 'edge_minmax', 'edge_zscore', 'edge_given_min_only'])
def test_normalize_equals_oracle(case):
    rng = np.random.default_rng(len(case))
    base = rng.normal(0, 5, (53, 17)) * rng.uniform(0.1, 100, 17)
    ref = rng.normal(0, 4, (40, 17)) * rng.uniform(0.1, 100, 17)
    _, (mins, maxs, _, _), _ = orc.py_normalize(ref)
    edge = _edge_matrix()
    if case == 'minmax':
        M, kw, flags = base, {}, orc.NONREAL
    elif case == 'zscore':
        M, kw, flags = base, dict(zscore=True), orc.NONREAL | orc.ZSCORE
    elif case == 'given_ranges':
        M, kw, flags = base, dict(mins=mins, maxs=maxs), orc.NONREAL | orc.G_MINS | orc.G_MAXS
    elif case == 'given_ranges_no_check':
        M, kw, flags = base, dict(mins=mins, maxs=maxs, non_real_check=False), orc.G_MINS | orc.G_MAXS
    elif case == 'edge_minmax':
        M, kw, flags = edge, {}, orc.NONREAL
    elif case == 'edge_zscore':
        M, kw, flags = edge, dict(zscore=True), orc.NONREAL | orc.ZSCORE
    elif case == 'edge_given_min_only':
        em = np.full(edge.shape[1], -1.0)
        M, kw, flags = edge, dict(mins=em), orc.NONREAL | orc.G_MINS
        mins = em
    else:
        M = edge
    print('normalize', case, flush=True)
    if case.startswith('replace'):
        g = M.copy()
        if case == 'replace_given':
            r = ReplaceNonReal(g, mins[:11], maxs[:11])
            want, _, rw = orc.py_normalize(M, flags=orc.REPLACE_ONLY | orc.G_MINS | orc.G_MAXS, mins=mins[:11],
                                           maxs=maxs[:11])
        else:
            r = ReplaceNonReal(g)
            want, _, rw = orc.py_normalize(M, flags=orc.REPLACE_ONLY)
        assert r == bool(rw)
        assert np.array_equal(g, want, equal_nan=True)
        return
    g = M.copy()
    out = normalize_by_columns(g, **kw)
    want, arrs, _ = orc.py_normalize(M, flags=flags, mins=kw.get('mins'), maxs=kw.get('maxs'))
    assert np.array_equal(g, want, equal_nan=True)
    if flags & orc.ZSCORE:
        assert np.array_equal(out[2], arrs[2], equal_nan=True) and np.array_equal(out[3], arrs[3], equal_nan=True)
    else:
        assert np.array_equal(out[0], arrs[0], equal_nan=True) and np.array_equal(out[1], arrs[1], equal_nan=True)


def _space(seed, sizes, nf):
    rng = np.random.default_rng(seed)
    centres = rng.normal(0, 1, (len(sizes), nf)) * rng.uniform(0.2, 4, nf)
    X = np.concatenate([centres[c] + rng.normal(0, 1, (n, nf)) for c, n in enumerate(sizes)])
    X[:, 2] = 7.0  # constant feature: weight 0
    names = ['f%d' % k for k in range(nf)]
    fs = FeatureSpace(name='synthetic', feature_names=names, data_matrix=X, class_names=['c%d' % (c + 1) for c in
                      range(len(sizes))], class_sizes=sizes,
                      sample_names=['s%d' % i for i in range(len(X))])
    return fs


@pytest.mark.parametrize('sizes', [[9, 6, 8], [40, 40], [1, 12], [5, 5, 5, 5, 5, 5]])
def test_fisher_equals_oracle(sizes):
    fs = _space(sum(sizes), sizes, 33).Normalize(inplace=True)
    got = np.array(FisherFeatureWeights.NewFromFeatureSpace(fs).values)
    w = np.zeros(33)
    cls = np.ascontiguousarray(fs._class_indices())
    M = np.ascontiguousarray(fs.data_matrix)
    orc.lib().orc_py_fisher(orc.ptr(M), orc.ptr(cls), M.shape[0], 33, len(sizes), orc.ptr(w))
    assert np.array_equal(got, w)
    assert got[2] == 0
    if min(sizes) < 2:
        assert np.all(got == 0)


@pytest.mark.parametrize('seed', [1, 2, 3])
def test_wnd5_equals_oracle(seed):
    sizes = [30, 25, 35, 10][:seed + 1]
    tr = _space(seed, sizes, 60).Normalize(inplace=True)
    rng = np.random.default_rng(seed + 100)
    te_m = tr.data_matrix[rng.integers(0, tr.num_samples, 12)] + rng.normal(0, 3, (12, 60))
    te_m[0] = tr.data_matrix[4]  # a collision with one training sample
    te = FeatureSpace(name='test', feature_names=tr.feature_names, data_matrix=te_m,
                      sample_names=['t%d' % i for i in range(12)])
    fw = FisherFeatureWeights.NewFromFeatureSpace(tr).Threshold(30)
    trr, ter = tr.FeatureReduce(fw), te.FeatureReduce(fw)
    res = FeatureSpaceClassification.NewWND5(trr, ter, fw).individual_results
    cls = np.ascontiguousarray(trr._class_indices())
    for q, r in enumerate(res):
        pred, prob, nf = orc.py_wnd5(trr.data_matrix, cls, len(sizes), np.array(fw.values), ter.data_matrix[q])
        assert r.predicted_class == pred
        np.testing.assert_allclose(r.marginal_probabilities, prob, rtol=1e-12)
        np.testing.assert_allclose(r.normalization_factor, nf, rtol=1e-12)


def test_wnd5_no_prediction_when_a_class_collides_entirely():
    tr = _space(9, [6, 6], 20).Normalize(inplace=True)
    tr.data_matrix[6:] = tr.data_matrix[0]  # class 2 = copies of sample 0
    te = FeatureSpace(name='test', feature_names=tr.feature_names, data_matrix=tr.data_matrix[:1].copy(),
                      sample_names=['t0'])
    r = FeatureSpaceClassification.NewWND5(tr, te, None).individual_results[0]
    assert r.marginal_probabilities is None and r.predicted_label is None


def test_generate_features_matches_sig_golden(test_dir):
    """FeatureVector.GenerateFeatures on the reference's t1 TIFF through the
    engine, against the reference's precalculated .sig"""
    fv = FeatureVector(source_filepath=os.path.join(test_dir, 't1_s01_c05_ij.tif'), long=True).GenerateFeatures()
    gold = FeatureVector(long=True).LoadSigFile(os.path.join(test_dir, 't1_s01_c05_ij-l_precalculated.sig'))
    assert sorted(fv.feature_names) == sorted(gold.feature_names)
    fv = fv.FeatureReduce(gold)
    bad = [(n, a, b) for n, a, b in zip(gold.feature_names, fv.values, gold.values)
           if not (n.startswith('Haralick Textures') and (n.endswith('[16]') or n.endswith('[17]')))
           and not wndcharm_compare(a, b)]
    assert not bad, bad[:5]
