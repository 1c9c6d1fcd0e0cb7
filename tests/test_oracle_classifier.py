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
