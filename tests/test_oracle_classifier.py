"""Pins the oracle's classifier (TrainingSet::normalize/SetFisherScores/
classify2, signatures::normalize) against the reference's known answers:
  * Fisher weights of test-l.fit (`wndchrm classify -l -f1.0 -v...`):
    tests/wndchrm_tests/test_fit-l.weights, delta 1e-5 as in
    tests/pywndcharm_tests/test_FeatureWeights.py:51-62;
  * the expected `wndchrm classify -l test-l.fit test-l.fit` table of
    tests/wndchrm_tests/test01-classify.pl:104-125 (norm factor to 3
    significant digits, marginal probabilities to 3 decimals, predicted class,
    interpolated value)."""
import numpy as np

from helpers import load_npz, oracle, ptr


def normalized(d):
    data = np.ascontiguousarray(d['values'].copy())
    n, f = data.shape
    mins, maxs = np.zeros(f), np.zeros(f)
    oracle().orc_normalize_train(ptr(data), n, f, ptr(mins), ptr(maxs))
    return data, mins, maxs


def test_normalize_closed_form():
    d = load_npz('test_l_fit.npz')
    data, mins, maxs = normalized(d)
    raw = d['values']
    rng = raw.max(0) - raw.min(0)
    ok = rng >= np.finfo(float).eps
    want = np.where(ok, 100 * ((raw - raw.min(0)) / np.where(ok, rng, 1)), 0)
    assert np.array_equal(data, want)
    assert data.min() >= 0 and data.max() <= 100


def test_fisher_weights_f1():
    d = load_npz('test_l_fit.npz')
    data, _, _ = normalized(d)
    n, f = data.shape
    w = np.zeros(f)
    cls = np.ascontiguousarray(d['classes'])
    oracle().orc_fisher(ptr(data), ptr(cls), n, f, 2, 1.0, ptr(w))
    np.testing.assert_allclose(w, d['weights_f1'], atol=1e-5 + 5e-7)


def test_wnd5_classify_table():
    d = load_npz('test_l_fit.npz')
    data, mins, maxs = normalized(d)
    n, f = data.shape
    cls = np.ascontiguousarray(d['classes'])
    w = np.zeros(f)
    oracle().orc_fisher(ptr(data), ptr(cls), n, f, 2, 0.15, ptr(w))
    assert (w > 0).sum() == 439  # threshold = sorted[floor(0.15*2919+0.5)] = sorted[438], kept inclusively
    for i in range(n):
        t = np.ascontiguousarray(d['values'][i].copy())
        oracle().orc_normalize_test(ptr(t), f, ptr(mins), ptr(maxs))
        prob = np.zeros(2)
        nf = np.zeros(1)
        pred = oracle().orc_wnd5(ptr(data), ptr(cls), n, f, 2, ptr(w), ptr(t), ptr(prob), ptr(nf))
        assert pred == d['kat_pred'][i]
        assert float('%.3g' % nf[0]) == d['kat_norm'][i]
        assert np.array_equal(np.array([float('%.3f' % p) for p in prob]), d['kat_p'][i])
        interp = prob[0] * 2 + prob[1] * 4
        assert float('%.3f' % interp) == d['kat_interp'][i]


def _wnn(train, cls, C, w, t):
    prob, nf = np.zeros(C), np.zeros(1)
    tr = np.ascontiguousarray(train)
    c = np.ascontiguousarray(cls, np.int32)
    pred = oracle().orc_wnn(ptr(tr), ptr(c), tr.shape[0], tr.shape[1], C, ptr(np.ascontiguousarray(w)),
                            ptr(np.ascontiguousarray(t)), ptr(prob), ptr(nf))
    return pred, prob, nf[0]


def test_wnn_is_weighted_minkowski_nearest_neighbour():
    """WNNclassify (TrainingSet.cpp:2375-2417): the distance is scipy's
    weighted Minkowski distance with p = 2 and the weights to the first power
    (TrainingSet::distance, 2357-2363); the class of the nearest sample is
    predicted; the marginals weight each class by the inverse of its closest
    distance, and norm_factor * marginal * distance = 1 as the reference's
    comment on WNNclassify states"""
    from scipy.spatial.distance import cdist
    rng = np.random.default_rng(41)
    n, f, C = 60, 14, 3
    train = rng.uniform(0, 100, (n, f))
    cls = np.repeat(np.arange(1, C + 1), n // C)
    w = rng.uniform(0, 2, f)
    w[3] = 0.0
    tests = rng.uniform(0, 100, (12, f))
    d = cdist(tests, train, 'minkowski', p=2, w=w)
    for k, t in enumerate(tests):
        pred, prob, nf = _wnn(train, cls, C, w, t)
        assert pred == cls[np.argmin(d[k])]
        dmin = np.array([d[k][cls == c].min() for c in range(1, C + 1)])
        np.testing.assert_allclose(prob[:, None] / prob[None, :], dmin[None, :] / dmin[:, None], rtol=1e-12)
        np.testing.assert_allclose(prob.sum(), 1.0, rtol=1e-14)
        np.testing.assert_allclose(nf * prob * dmin, 1.0, rtol=1e-12)
        assert np.argmax(prob) + 1 == pred


def test_wnn_identical_sample_and_empty_class():
    """a training vector equal to the test vector is skipped (distance below
    1/INF counts as INF), the first of two equally close samples wins, and a
    class with no sample keeps INF (its marginal 1e-201 / sum)"""
    train = np.array([[0.0, 0.0], [3.0, 4.0], [1.0, 1.0], [-1.0, -1.0]])
    cls = np.array([1, 2, 2, 1])
    w = np.ones(2)
    pred, prob, nf = _wnn(train, cls, 3, w, np.array([0.0, 0.0]))  # sample 0 is skipped; 2 and 3 tie
    assert pred == 2
    assert prob[0] == prob[1] and prob[2] > 0 and prob[2] < 1e-200
    pred, prob, nf = _wnn(train, cls, 3, w, np.array([3.0, 4.0]))  # nearest other: (1, 1), class 2
    assert pred == 2
    np.testing.assert_allclose(prob[1] / prob[0], np.hypot(3, 4) / np.hypot(2, 3), rtol=1e-14)
