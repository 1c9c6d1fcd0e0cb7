"""Pins the oracle's restatement of the wndcharm Python package's classifier
maths (orc_np_sum, orc_py_normalize, orc_py_fisher, orc_py_wnd5) to:

* numpy and scipy themselves, for the reductions the Python code delegates
  to them: numpy's 1-D sum, axis-0 mean / std, scipy's cdist
  'sqeuclidean', scipy.stats.mstats.zscore;
* the reference's own known answers:
  - FeatureSpace.Normalize of test-l.fit against test_fit-l-normalized.fit
    (tests/pywndcharm_tests/test_FeatureSpace.py:724-733, rtol 1e-5);
  - FisherFeatureWeights.NewFromFeatureSpace against test_fit-l.weights
    (tests/pywndcharm_tests/test_FeatureWeights.py:51-62, delta 1e-5);
  - SingleSampleClassification.NewWND5 marginal probabilities at 2189 / 438
    / 146 features (tests/pywndcharm_tests/test_SingleSamplePrediction.py:
    57-91, three decimals);
* closed forms for the non-real handling of ReplaceNonReal
  (wndcharm/utils.py:124-169)."""
import ctypes
import gzip
import os

import numpy as np
import pytest

from helpers import GOLDEN, oracle, ptr

P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double


def lib():
    L = oracle()
    L.orc_np_sum.argtypes = [P, ctypes.c_long]
    L.orc_np_sum.restype = D
    L.orc_py_normalize.argtypes = [P, I, I, I, P, P, P, P]
    L.orc_py_normalize.restype = I
    L.orc_py_fisher.argtypes = [P, P, I, I, I, P]
    L.orc_py_wnd5.argtypes = [P, P, I, I, I, P, P, P, P]
    L.orc_py_wnd5.restype = ctypes.c_long
    return L


ZSCORE, NONREAL, G_MINS, G_MAXS, G_MEANS, G_STDEVS, REPLACE_ONLY = 1, 2, 4, 8, 16, 32, 64


def py_normalize(M, flags=NONREAL, mins=None, maxs=None, means=None, stdevs=None):
    M = np.ascontiguousarray(M, np.float64).copy()
    n, f = M.shape
    arrs = [np.ascontiguousarray(a, np.float64).copy() if a is not None else np.zeros(f)
            for a in (mins, maxs, means, stdevs)]
    r = lib().orc_py_normalize(ptr(M), n, f, flags, *[ptr(a) for a in arrs])
    return M, arrs, r


def read_fit_text(name):
    """(values [n][f], classes, feature names) of a .fit fixture, parsed here
    with plain Python (TrainingSet::SaveToFile layout)."""
    lines = gzip.open(os.path.join(GOLDEN, 'raw', name), 'rt').read().splitlines()
    nf, ns = int(lines[1]), int(lines[2])
    names = lines[3:3 + nf]
    ncls = int(lines[0].split()[0])
    pos = 3 + nf + 1 + ncls
    vals, cls = [], []
    for s in range(ns):
        toks = lines[pos + 2 * s].split()
        vals.append([float(t) for t in toks[:nf]])
        cls.append(int(toks[nf]))
    return np.array(vals), np.array(cls, np.int32), names


def read_weights_text():
    rows = [l.split(None, 1) for l in gzip.open(os.path.join(GOLDEN, 'raw', 'test_fit-l.weights.gz'), 'rt')
            .read().splitlines() if l.strip()]
    return np.array([float(v) for v, _ in rows]), [n for _, n in rows]


def read_sig_text():
    lines = gzip.open(os.path.join(GOLDEN, 'raw', 't1_s01_c05_ij-l_precalculated.sig.gz'), 'rt').read().splitlines()
    vals, names = [], []
    for l in lines[2:]:
        v, n = l.split('\t', 1)
        vals.append(float(v))
        names.append(n)
    return np.array(vals), names


@pytest.mark.parametrize('n', [0, 1, 5, 7, 8, 9, 15, 16, 17, 127, 128, 129, 300, 1000, 4097, 8192, 8193, 20000])
def test_np_sum_is_numpy_sum(n):
    rng = np.random.default_rng(n)
    v = rng.normal(0, 1, n) * 10.0 ** rng.uniform(-8, 8, n)
    assert lib().orc_np_sum(ptr(v), n) == np.sum(v)


def test_min_max_normalize_reference_fit():
    raw, _, names = read_fit_text('test-l.fit.gz')
    target, _, tnames = read_fit_text('test_fit-l-normalized.fit.gz')
    assert names == tnames
    M, (mins, maxs, _, _), _ = py_normalize(raw)
    np.testing.assert_allclose(M, target, rtol=1e-5)
    # the ranges are numpy's axis-0 minima and maxima
    assert np.array_equal(mins, raw.min(0)) and np.array_equal(maxs, raw.max(0))


def test_zscore_normalize_matches_scipy():
    """z-scores against scipy's (the reference test's target).  On constant
    columns the two can differ by construction: the Python package divides
    by ndarray.std(axis=0), and where the axis-0 running mean of the 20
    equal values is off by an ulp (18 of the 410 constant columns of
    test-l.fit) the deviation is ~1e-20 rather than 0 and every z-score is
    exactly 1 (scipy's masked-array moments give 0).  Constant columns are
    held to numpy's own expression instead."""
    from scipy.stats.mstats import zscore
    raw, _, _ = read_fit_text('test-l.fit.gz')
    M, (_, _, means, stdevs), _ = py_normalize(raw, flags=NONREAL | ZSCORE)
    assert np.array_equal(means, raw.mean(0)) and np.array_equal(stdevs, raw.std(0))
    with np.errstate(all='ignore'):
        target = np.asarray(zscore(raw))
        expr = (raw - raw.mean(0)) / raw.std(0)
    target[np.isnan(target)] = 0
    expr[:, raw.std(0) == 0] = 0
    const = np.ptp(raw, axis=0) == 0
    np.testing.assert_allclose(M[:, ~const], target[:, ~const], rtol=1e-7, atol=1e-12)
    assert np.array_equal(M[:, const], expr[:, const])


def test_normalize_against_reference_ranges():
    """a test set normalised against a training set's ranges: values outside
    are clipped to the range, then scaled"""
    rng = np.random.default_rng(4)
    tr = rng.normal(0, 1, (30, 6))
    te = rng.normal(0, 2, (9, 6))
    _, (mins, maxs, _, _), _ = py_normalize(tr)
    M, _, r = py_normalize(te, flags=NONREAL | G_MINS | G_MAXS, mins=mins, maxs=maxs)
    assert r == 0
    assert M.min() >= 0 and M.max() <= 100
    inside = (te >= mins) & (te <= maxs)
    np.testing.assert_allclose(M[inside], (100 * (te - mins) / (maxs - mins))[inside], rtol=1e-14)
    assert np.all(M[te < mins] == 0) and np.all(M[te > maxs] == 100)


def test_replace_non_real_closed_forms():
    M = np.array([[1.0, np.nan, 5.0, np.inf],
                  [np.inf, 2.0, -1.0, np.nan],
                  [3.0, -np.inf, 7.0, np.nan]])
    # ReplaceNonReal alone, no bounds: +inf -> the finite column maximum, -inf
    # stays (only the upper bound is set), NaN -> 0; an all-non-finite
    # column clips +inf to the masked fill value 1e20
    R, _, r = py_normalize(M, flags=REPLACE_ONLY)
    assert r == 1
    want = np.array([[1.0, 0.0, 5.0, 1e20], [3.0, 2.0, -1.0, 0.0], [3.0, -np.inf, 7.0, 0.0]])
    assert np.array_equal(R, want)
    # nothing to replace: unchanged, returns 0
    F = np.arange(12.0).reshape(3, 4)
    R2, _, r2 = py_normalize(F, flags=REPLACE_ONLY)
    assert r2 == 0 and np.array_equal(R2, F)
    # with reference bounds: both bounds clip, NaN -> 0
    mins, maxs = np.array([0.0, 0.0, 0.0, 0.0]), np.array([2.0, 2.0, 2.0, 2.0])
    R3, _, r3 = py_normalize(M, flags=REPLACE_ONLY | G_MINS | G_MAXS, mins=mins, maxs=maxs)
    assert r3 == 1
    assert np.array_equal(R3, np.array([[1.0, 0.0, 2.0, 2.0], [2.0, 2.0, 0.0, 0.0], [2.0, 0.0, 2.0, 0.0]]))
    # a constant column normalises to 0
    C = np.array([[4.0, 1.0], [4.0, 3.0]])
    R4, (mn, mx, _, _), _ = py_normalize(C)
    assert np.array_equal(R4, np.array([[0.0, 0.0], [0.0, 100.0]]))


def test_fisher_reference_weights():
    raw, cls, _ = read_fit_text('test-l.fit.gz')
    M, _, _ = py_normalize(raw)
    w = np.zeros(M.shape[1])
    lib().orc_py_fisher(ptr(M), ptr(cls), M.shape[0], M.shape[1], 2, ptr(w))
    target, _ = read_weights_text()
    np.testing.assert_allclose(w, target, atol=1e-5 + 5e-7, rtol=0)


def test_fisher_matches_numpy_reductions_and_edge_cases():
    """three classes of unequal sizes; a class of one sample makes its
    variance NaN (0/0) and every weight 0; a constant feature has zero
    within-class variance -> NaN -> 0"""
    rng = np.random.default_rng(6)
    X = rng.normal(0, 1, (23, 5))
    X[:, 3] = 2.5
    cls = np.array([1] * 9 + [2] * 6 + [3] * 8, np.int32)
    w = np.zeros(5)
    lib().orc_py_fisher(ptr(X), ptr(cls), 23, 5, 3, ptr(w))
    pop = X.mean(0)
    mc = np.array([X[cls == c].mean(0) for c in (1, 2, 3)])
    vc = np.array([X[cls == c].var(0, ddof=1) for c in (1, 2, 3)])
    with np.errstate(all='ignore'):
        den = vc.mean(0)
        den[den == 0] = np.nan
        want = np.ma.masked_invalid((np.square(pop - mc).sum(0) / 2) / den).filled(0)
    assert np.array_equal(w, want)
    assert w[3] == 0 and np.all(w[[0, 1, 2, 4]] > 0)
    cls1 = np.array([1] * 22 + [2], np.int32)
    w1 = np.zeros(5)
    lib().orc_py_fisher(ptr(X), ptr(cls1), 23, 5, 2, ptr(w1))
    assert np.all(w1 == 0)


def py_wnd5(train, cls, C, w, test):
    prob, nf = np.zeros(C), np.zeros(1)
    # contiguous copies held by name for the duration of the call
    tr = np.ascontiguousarray(train, np.float64)
    cl = np.ascontiguousarray(cls, np.int32)
    wt = np.ascontiguousarray(w, np.float64)
    te = np.ascontiguousarray(test, np.float64)
    pred = lib().orc_py_wnd5(ptr(tr), ptr(cl), tr.shape[0], tr.shape[1], C, ptr(wt), ptr(te), ptr(prob), ptr(nf))
    return pred, prob, nf[0]


def test_wnd5_matches_scipy_and_numpy():
    from scipy.spatial.distance import cdist
    rng = np.random.default_rng(7)
    tr = rng.uniform(0, 100, (40, 30))
    cls = np.repeat([1, 2, 3, 4], 10).astype(np.int32)
    w = rng.uniform(0, 2, 30)
    w[:4] = 0
    te = rng.uniform(0, 100, 30)
    pred, prob, nf = py_wnd5(tr, cls, 4, w, te)
    d = cdist(tr * w, (te * w)[None, :], 'sqeuclidean')[:, 0]
    sims = np.power(np.ma.masked_less_equal(d, np.finfo(float).eps), -5)
    means = np.array([sims[cls == c].compressed().mean() for c in (1, 2, 3, 4)])
    np.testing.assert_allclose(prob, means / means.sum(), rtol=1e-13)
    np.testing.assert_allclose(nf, means.sum(), rtol=1e-13)
    assert pred == np.argmax(means) + 1
    # an exact training copy collides (distance 0, masked), not infinite
    te2 = tr[3].copy()
    pred2, prob2, _ = py_wnd5(tr, cls, 4, w, te2)
    assert np.all(np.isfinite(prob2))
    # a class whose every sample collides has no mean: no prediction
    tr3 = tr.copy()
    tr3[cls == 2] = te
    pred3, prob3, nf3 = py_wnd5(tr3, cls, 4, w, te)
    assert pred3 == 0 and np.all(np.isnan(prob3)) and np.isnan(nf3)


def _top(values, names, n):
    """FisherFeatureWeights.Threshold (FeatureWeights.py:250-306): the top n
    by value, descending, ties in their file order"""
    order = np.argsort(-np.asarray(values), kind='stable')[:n]
    return np.asarray(values)[order], [names[i] for i in order]


@pytest.mark.parametrize('nfeat,want', [(2189, [0.083, 0.917]), (438, [0.076, 0.924]), (146, [0.066, 0.934])])
def test_single_sample_wnd5_known_answers(nfeat, want):
    raw, cls, names = read_fit_text('test-l.fit.gz')
    M, (mins, maxs, _, _), _ = py_normalize(raw)
    sv, snames = read_sig_text()
    assert snames == names
    V, _, _ = py_normalize(sv[None, :], flags=NONREAL | G_MINS | G_MAXS, mins=mins, maxs=maxs)
    wv, wn = read_weights_text()
    assert wn == names
    tv, tn = _top(wv, wn, nfeat)
    idx = [names.index(n) for n in tn]
    pred, prob, _ = py_wnd5(M[:, idx], cls, 2, tv, V[0, idx])
    assert [round(p, 3) for p in prob] == want
    assert pred == 2  # 4cell
