"""Pins the FP64 torch reference of the differentiable path (tests/ad_reference.py)
to the oracle: forward values against the oracle's transforms, Zernike
features and WND5 probabilities; gradients (torch autograd) against central
finite differences of the oracle's own forward pass.  CPU only."""
import numpy as np
import pytest
import torch

import ad_reference as R
from helpers import (T_CHEBYSHEV, T_FOURIER, oracle, ptr)

ZERNIKE = 2  # ORC_A_ZERNIKE


def orc_transform(kind, img):
    img = np.ascontiguousarray(img, np.float64)
    h, w = img.shape
    out = np.zeros(max(h, w) * max(h, w))
    oh, ow = np.zeros(1, np.int32), np.zeros(1, np.int32)
    assert oracle().orc_transform(kind, ptr(img), h, w, ptr(out), ptr(oh), ptr(ow)) == 0
    return out[:oh[0] * ow[0]].reshape(oh[0], ow[0])


def orc_zernike(img):
    img = np.ascontiguousarray(img, np.float64)
    out = np.zeros(72)
    assert oracle().orc_algorithm(ZERNIKE, ptr(img), img.shape[0], img.shape[1], ptr(out)) == 0
    return out


def orc_cf(img):
    img = np.ascontiguousarray(img, np.float64)
    out = np.zeros(529)
    nn = np.zeros(1, np.int32)
    assert oracle().orc_cf_magnitudes(ptr(img), img.shape[0], img.shape[1], ptr(out), ptr(nn)) == 0
    return out[:nn[0] * nn[0]]


def orc_comb(img):
    img = np.ascontiguousarray(img, np.float64)
    out = np.zeros(320)
    assert oracle().orc_comb_raw(ptr(img), img.shape[0], img.shape[1], ptr(out)) == 0
    return out


def orc_gabor(img):
    img = np.ascontiguousarray(img, np.float64)
    out = np.zeros((8,) + img.shape)
    assert oracle().orc_gabor_energy(ptr(img), img.shape[0], img.shape[1], ptr(out)) == 0
    return out


def orc_node(img, chain):
    for t in chain:
        img = orc_transform({'Fourier': T_FOURIER, 'Chebyshev': T_CHEBYSHEV}[t], img)
    return img


def orc_features(img, groups):
    out = []
    for alg, chain in groups:
        p = orc_node(img, chain)
        if alg == 'Zernike Coefficients':
            out.append(orc_zernike(p))
        elif alg == 'Chebyshev-Fourier Coefficients (raw)':
            out.append(orc_cf(p))
        elif alg == 'Comb Moments (raw)':
            out.append(orc_comb(p))
        elif alg == 'Gabor Textures (raw)':
            out.append(orc_gabor(p).mean((1, 2)))
        else:  # raw order-20 Chebyshev coefficients = leading block of the full transform
            c = orc_transform(T_CHEBYSHEV, p)
            n = min(20, *p.shape)
            out.append(c[:n, :n].reshape(-1))
    return np.concatenate(out)


def plane(seed, h, w):
    from wndcharm import synthetic
    return synthetic.image(seed, 8, (h, w)).astype(np.float64) / 4095.0


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_fourier_reference_matches_oracle():
    img = plane(11, 40, 48)
    got = R.fourier(torch.from_numpy(img)).numpy()
    assert rel(got, orc_transform(T_FOURIER, img)) < 1e-13


def test_chebyshev_reference_matches_oracle():
    img = plane(12, 40, 48)
    want = orc_transform(T_CHEBYSHEV, img)
    got = R.chebyshev(torch.from_numpy(img)).numpy()
    assert got.shape == want.shape == (40, 40)
    assert rel(got, want) < 1e-12
    # each coefficient is independent of the order: the order-20 block is the leading block
    got20 = R.chebyshev(torch.from_numpy(img), 20).numpy()
    assert rel(got20, want[:20, :20]) < 1e-12


@pytest.mark.parametrize('shape', [(40, 48), (33, 29)])
def test_zernike_reference_matches_oracle(shape):
    img = plane(13, *shape)
    want = orc_zernike(img)
    got = R.zernike(torch.from_numpy(img)).numpy()
    assert np.all(np.abs(got - want) <= 1e-10 * np.abs(want) + 1e-14)


@pytest.mark.parametrize('shape', [(40, 48), (33, 29), (301, 300), (320, 352)])
def test_cf_reference_matches_oracle(shape):
    """raw Chebyshev-Fourier magnitudes, with the Downsample above 300 x 300"""
    img = plane(14, *shape)
    want = orc_cf(img)
    got = R.cf_magnitudes(torch.from_numpy(img)).numpy()
    assert got.shape == want.shape
    assert rel(got, want) < 1e-12


@pytest.mark.parametrize('shape', [(40, 48), (33, 29), (64, 20), (20, 64)])
def test_comb_reference_matches_oracle(shape):
    """raw comb line moments (two-pass here, Welford in the oracle)"""
    img = plane(17, *shape)
    want = orc_comb(img)
    got = R.comb_raw(torch.from_numpy(img)).numpy()
    assert np.all(np.abs(got - want) <= 1e-9 * np.abs(want) + 1e-12)


@pytest.mark.parametrize('shape', [(40, 48), (20, 64)])
def test_gabor_reference_matches_oracle(shape):
    """energy maps: the oracle's conv2comp order against torch's correlation"""
    img = plane(19, *shape)
    want = orc_gabor(img)
    got = R.gabor_energy(torch.from_numpy(img)).numpy()
    assert rel(got, want) < 1e-12


def _training_set(n_classes, per_class, groups, shape, seed):
    lib = oracle()
    feats, cls = [], []
    for c in range(n_classes):
        for k in range(per_class):
            from wndcharm import synthetic
            img = synthetic.image(seed + 100 * c + k, [6, 12, 24][c % 3], shape).astype(np.float64) / 4095.0
            feats.append(orc_features(img, groups))
            cls.append(c + 1)
    data = np.ascontiguousarray(feats, np.float64)
    n, f = data.shape
    mins, maxs = np.zeros(f), np.zeros(f)
    lib.orc_normalize_train(ptr(data), n, f, ptr(mins), ptr(maxs))
    cls = np.asarray(cls, np.int32)
    w = np.zeros(f)
    lib.orc_fisher(ptr(data), ptr(cls), n, f, n_classes, 0.5, ptr(w))
    return data, cls, w, mins, maxs


def orc_probabilities(x, data, cls, w, mins, maxs, n_classes):
    v = np.ascontiguousarray(x, np.float64).copy()
    lib = oracle()
    lib.orc_normalize_test(ptr(v), len(v), ptr(mins), ptr(maxs))
    prob = np.zeros(n_classes)
    norm = np.zeros(1)
    lib.orc_wnd5(ptr(data), ptr(cls), data.shape[0], data.shape[1], n_classes, ptr(w), ptr(v), ptr(prob), ptr(norm))
    return prob


GROUPS = [('Zernike Coefficients', []), ('Zernike Coefficients', ['Fourier']),
          ('Zernike Coefficients', ['Chebyshev']), ('Chebyshev Coefficients (raw)', ['Fourier'])]


def test_wnd_reference_matches_oracle():
    shape = (24, 28)
    data, cls, w, mins, maxs = _training_set(3, 4, GROUPS, shape, 500)
    x = np.stack([orc_features(plane(900 + k, *shape), GROUPS) for k in range(3)])
    want = np.stack([orc_probabilities(v, data, cls, w, mins, maxs, 3) for v in x])
    t = torch.from_numpy
    got = R.wnd5(t(x), t(data), t(cls), t(w), t(mins), t(maxs), 3).numpy()
    assert rel(got, want) < 1e-10


@pytest.mark.parametrize('group', GROUPS + [('Zernike Coefficients', ['Chebyshev', 'Fourier']),
                                   ('Chebyshev-Fourier Coefficients (raw)', []),
                                   ('Chebyshev-Fourier Coefficients (raw)', ['Fourier']),
                                   ('Comb Moments (raw)', []), ('Comb Moments (raw)', ['Chebyshev']),
                                   ('Gabor Textures (raw)', [])],
                         ids=lambda g: g[0].split()[0] + '(' + ','.join(g[1]) + ')')
def test_reference_gradient_matches_oracle_differences(group):
    """d(c . features)/dI from torch autograd = central differences of the oracle"""
    shape = (24, 28)
    img = plane(21, *shape)
    rng = np.random.default_rng(5)
    nf = len(orc_features(img, [group]))
    c = rng.standard_normal(nf)
    x = torch.from_numpy(img.copy()).requires_grad_(True)
    (R.features(x, [group]) * torch.from_numpy(c)).sum().backward()
    g = x.grad.numpy()
    h = 1e-4
    for idx in rng.choice(img.size, 12, replace=False):
        p = np.unravel_index(idx, shape)
        a, b = img.copy(), img.copy()
        a[p] += h
        b[p] -= h
        fd = (c @ orc_features(a, [group]) - c @ orc_features(b, [group])) / (2 * h)
        assert abs(fd - g[p]) <= 1e-5 * np.max(np.abs(g)), (p, fd, g[p])


def test_reference_wnd_gradient_matches_oracle_differences():
    shape = (24, 28)
    data, cls, w, mins, maxs = _training_set(3, 4, GROUPS, shape, 700)
    img = plane(31, *shape)
    t = torch.from_numpy
    x = t(img.copy()).requires_grad_(True)
    p = R.wnd5(R.features(x, GROUPS)[None], t(data), t(cls), t(w), t(mins), t(maxs), 3)[0]
    cls_k = int(torch.argmin(p))  # a class whose probability is not saturated at 1
    p[cls_k].backward()
    g = x.grad.numpy()
    assert np.max(np.abs(g)) > 0
    rng = np.random.default_rng(9)
    h = 1e-5
    for idx in rng.choice(img.size, 12, replace=False):
        q = np.unravel_index(idx, shape)
        a, b = img.copy(), img.copy()
        a[q] += h
        b[q] -= h
        pa = orc_probabilities(orc_features(a, GROUPS), data, cls, w, mins, maxs, 3)[cls_k]
        pb = orc_probabilities(orc_features(b, GROUPS), data, cls, w, mins, maxs, 3)[cls_k]
        fd = (pa - pb) / (2 * h)
        assert abs(fd - g[q]) <= 1e-4 * np.max(np.abs(g)), (q, fd, g[q])


def test_wnd_model_rejects_mismatched_training_set():
    """feature rows and class labels must agree before anything reaches the GPU"""
    import wndcharm
    from wndcharm import autograd
    with pytest.raises(wndcharm.WndGpuError):
        autograd.WNDModel(np.zeros((10, 4)), np.ones(8, np.int32), 2)
    with pytest.raises(wndcharm.WndGpuError):
        autograd.WNDModel(np.zeros((3, 4)), np.array([1, 2, 3], np.int32), 2)
    with pytest.raises(wndcharm.WndGpuError):
        wndcharm.fisher_weights(np.zeros((10, 4)), np.ones(9, np.int32), 2)
    with pytest.raises(wndcharm.WndGpuError):
        wndcharm.wnd5(np.zeros((10, 4)), np.ones(10, np.int32), 2, np.ones(4), np.zeros((2, 5)))
