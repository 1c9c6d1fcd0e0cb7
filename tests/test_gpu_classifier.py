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


def test_short_set_five_fold_predictions_match_oracle():
    """The plain-method short-set configuration: 2 classes x 10 seeded 128x128
    images, short-set features on the GPU against the oracle (integer-valued
    columns exact, the rest to 1e-6), then a 5-fold train/test split (two
    images per class held out per fold) with normalisation, Fisher weights
    (top 15%) and WND5, whose predicted classes must equal the oracle's run
    on the oracle's features."""
    from wndcharm import synthetic
    from helpers import exact_mask_for_names, orc_features, orc_names
    imgs, cls = synthetic.config_images('short_128')
    plan = wndcharm.StdFeatureComputationPlans.getFeatureSet()
    got = wndcharm.compute_features(imgs, plan)
    want = np.stack([orc_features(im.astype(np.float64), False) for im in imgs])
    exact = exact_mask_for_names(orc_names(False))
    assert np.array_equal(got[:, exact], want[:, exact], equal_nan=True)
    close = (got == want) | (np.abs(got - want) <= 1e-6 * np.abs(want)) | (np.abs(got - want) < 1e-12)
    assert close[:, ~exact].all()
    lib = oracle()
    for k in range(5):
        test = np.r_[2 * k:2 * k + 2, 10 + 2 * k:10 + 2 * k + 2]
        train = np.setdiff1d(np.arange(20), test)
        tr, mins, maxs = wndcharm.normalize_train(got[train])
        w = wndcharm.fisher_weights(tr, cls[train], 2, 0.15)
        te = np.stack([wndcharm.normalize_test(got[t], mins, maxs) for t in test])
        pred = wndcharm.wnd5(tr, cls[train], 2, w, te)[2]
        otr = np.ascontiguousarray(want[train])
        n, f = otr.shape
        omin, omax = np.zeros(f), np.zeros(f)
        lib.orc_normalize_train(ptr(otr), n, f, ptr(omin), ptr(omax))
        ocls = np.ascontiguousarray(cls[train], np.int32)
        ow = np.zeros(f)
        lib.orc_fisher(ptr(otr), ptr(ocls), n, f, 2, 0.15, ptr(ow))
        for j, t in enumerate(test):
            v = np.ascontiguousarray(want[t].copy())
            lib.orc_normalize_test(ptr(v), f, ptr(omin), ptr(omax))
            prob, norm = np.zeros(2), np.zeros(1)
            opred = lib.orc_wnd5(ptr(otr), ptr(ocls), n, f, 2, ptr(ow), ptr(v), ptr(prob), ptr(norm))
            assert pred[j] == opred, (k, t)


def test_wnn_matches_oracle():
    """WNNclassify on the GPU (wndgpu_wnn) against the oracle's: predicted
    classes equal, probabilities and norm factors within 1e-12 (the GPU takes
    d*d and sqrt where the reference calls glibc's pow), test vectors that
    duplicate a training vector included"""
    rng = np.random.default_rng(77)
    n, f, C = 240, 300, 4
    train = rng.uniform(0, 100, (n, f))
    cls = np.repeat(np.arange(1, C + 1), n // C).astype(np.int32)
    w = rng.uniform(0, 1, f)
    w[rng.random(f) < 0.3] = 0.0
    tests = rng.uniform(0, 100, (40, f))
    tests[::7] = train[::35][:len(tests[::7])]
    print('wnn', n, f, C, len(tests), flush=True)
    prob, norm, pred = wndcharm.wnn(train, cls, C, w, tests)
    lib = oracle()
    for k, t in enumerate(tests):
        op, on = np.zeros(C), np.zeros(1)
        opred = lib.orc_wnn(ptr(np.ascontiguousarray(train)), ptr(cls), n, f, C, ptr(np.ascontiguousarray(w)),
                            ptr(np.ascontiguousarray(t)), ptr(op), ptr(on))
        assert pred[k] == opred, k
        np.testing.assert_allclose(prob[k], op, rtol=1e-12)
        np.testing.assert_allclose(norm[k], on[0], rtol=1e-12)
