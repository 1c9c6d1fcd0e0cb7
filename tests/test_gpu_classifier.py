"""GPU classifier (normalise, Fisher, WND5) against the oracle and the
reference's classify known answers (tests/wndchrm_tests/test01-classify.pl)."""
import numpy as np
import pytest

import wndcharm
from helpers import load_npz, oracle, ptr

pytestmark = pytest.mark.gpu


def test_normalize_fisher_wnd5():
    d = load_npz('test_l_fit.npz')
    raw = d['values']
    cls = np.ascontiguousarray(d['classes'])
    n, f = raw.shape
    norm, mins, maxs = wndcharm.normalize_train(raw)
    ref = np.ascontiguousarray(raw.copy())
    rmin, rmax = np.zeros(f), np.zeros(f)
    oracle().orc_normalize_train(ptr(ref), n, f, ptr(rmin), ptr(rmax))
    assert np.array_equal(norm, ref) and np.array_equal(mins, rmin) and np.array_equal(maxs, rmax)
    for used in (1.0, 0.15):
        w = wndcharm.fisher_weights(norm, cls, 2, used)
        rw = np.zeros(f)
        oracle().orc_fisher(ptr(ref), ptr(cls), n, f, 2, used, ptr(rw))
        assert np.array_equal(w, rw)
    np.testing.assert_allclose(wndcharm.fisher_weights(norm, cls, 2, 1.0), d['weights_f1'], atol=1e-5 + 5e-7)
    w = wndcharm.fisher_weights(norm, cls, 2, 0.15)
    tests = np.array([wndcharm.normalize_test(raw[i], mins, maxs) for i in range(n)])
    prob, nf, pred = wndcharm.wnd5(norm, cls, 2, w, tests)
    assert list(pred) == list(d['kat_pred'])
    for i in range(n):
        assert float('%.3g' % nf[i]) == d['kat_norm'][i]
        assert [float('%.3f' % p) for p in prob[i]] == list(d['kat_p'][i])
