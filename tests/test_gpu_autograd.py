"""Differentiable form on the GPU: forward values of the adjoint path against the
oracle, adjoint kernels against the FP64 torch reference gradient (pinned to
the oracle by tests/test_ad_reference.py), and against central finite
differences of the GPU forward pass (h = 1e-3, 256 random pixels) through the
whole chain pixels -> features -> WND5 probability."""
import numpy as np
import pytest
import torch

import ad_reference as R
from test_ad_reference import (GROUPS, orc_cf, orc_comb, orc_features, orc_gabor, orc_probabilities,
                               orc_transform, orc_zernike, plane)
from helpers import T_CHEBYSHEV, T_FOURIER

pytestmark = pytest.mark.gpu

DEV = torch.device('cuda', 0)


def _ad():
    from wndcharm import autograd
    return autograd


def batch(seeds, h, w):
    return torch.from_numpy(np.stack([plane(s, h, w) for s in seeds])).to(DEV)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize('shape', [(40, 48), (64, 64), (45, 37)])
def test_forward_matches_oracle(shape):
    A = _ad()
    x = batch([1, 2, 3], *shape)
    xs = x.cpu().numpy()
    f = A.op_forward(A.FOURIER, x).cpu().numpy()
    c = A.op_forward(A.CHEBYSHEV, x).cpu().numpy()
    c20 = A.op_forward(A.CHEBYSHEV_COEFFS, x).cpu().numpy()
    z = A.op_forward(A.ZERNIKE, x).cpu().numpy().reshape(3, 72)
    for i in range(3):
        assert rel(f[i], orc_transform(T_FOURIER, xs[i])) < 1e-12
        want_c = orc_transform(T_CHEBYSHEV, xs[i])
        assert rel(c[i], want_c) < 1e-12
        n = min(20, *shape)
        assert rel(c20[i], want_c[:n, :n]) < 1e-12
        wz = orc_zernike(xs[i])
        assert np.all(np.abs(z[i] - wz) <= 1e-9 * np.abs(wz) + 1e-14)


@pytest.mark.parametrize('shape', [(64, 64), (48, 56), (301, 300), (320, 352)])
def test_cheb_fourier_forward_matches_oracle(shape):
    """raw Chebyshev-Fourier magnitudes, with the Downsample above 300 x 300"""
    A = _ad()
    x = batch([15, 16], *shape)
    got = A.cheb_fourier(x).cpu().numpy()
    for i in range(2):
        want = orc_cf(x[i].cpu().numpy())
        assert got[i].shape == want.shape
        assert rel(got[i], want) < 1e-11


@pytest.mark.parametrize('shape', [(64, 64), (40, 48), (33, 29), (512, 512)])
def test_comb_moments_forward_matches_oracle(shape):
    A = _ad()
    x = batch([18, 19], *shape)
    got = A.comb_moments(x).cpu().numpy()
    for i in range(2):
        want = orc_comb(x[i].cpu().numpy())
        assert np.all(np.abs(got[i] - want) <= 1e-9 * np.abs(want) + 1e-12)


@pytest.mark.parametrize('shape', [(64, 64), (40, 100), (33, 29)])
def test_gabor_energy_forward_matches_oracle(shape):
    A = _ad()
    x = batch([20, 21], *shape)
    got = A.gabor_energy(x).cpu().numpy()
    for i in range(2):
        assert rel(got[i], orc_gabor(x[i].cpu().numpy())) < 1e-12


def test_zernike_matches_feature_path():
    import wndcharm
    A = _ad()
    x = batch([4, 5], 64, 80)
    z = A.zernike(x).cpu().numpy()
    alg = wndcharm.FeatureAlgorithm('Zernike Coefficients')
    for i in range(2):
        want = alg.execute(x[i].cpu().numpy())
        assert np.all(np.abs(z[i] - want) <= 1e-9 * np.abs(want) + 1e-14)


@pytest.mark.parametrize('op', ['fourier', 'chebyshev', 'chebyshev_coeffs', 'zernike', 'cheb_fourier',
                                'comb_moments', 'gabor_energy'])
@pytest.mark.parametrize('shape', [(40, 48), (64, 64), (310, 320)])
def test_adjoint_matches_reference_gradient(op, shape):
    A = _ad()
    x = batch([7, 8], *shape)
    fn = getattr(A, op)
    ref = {'fourier': R.fourier, 'chebyshev': R.chebyshev, 'chebyshev_coeffs': lambda p: R.chebyshev(p, 20),
           'zernike': R.zernike, 'cheb_fourier': R.cf_magnitudes, 'comb_moments': R.comb_raw,
           'gabor_energy': R.gabor_energy}[op]
    xg = x.clone().requires_grad_(True)
    y = fn(xg)
    rng = np.random.default_rng(3)
    g = torch.from_numpy(rng.standard_normal(tuple(y.shape))).to(DEV)
    (y * g).sum().backward()
    got = xg.grad.cpu().numpy()
    for i in range(2):
        xr = x[i].cpu().clone().requires_grad_(True)
        (ref(xr).reshape(-1) * g[i].cpu().reshape(-1)).sum().backward()
        assert rel(got[i], xr.grad.numpy()) < 1e-9


NAMES = ['Zernike Coefficients ()', 'Zernike Coefficients (Fourier ())', 'Zernike Coefficients (Chebyshev ())',
         'Chebyshev Coefficients (raw) (Fourier ())']


def test_plan_features_and_gradient_match_reference():
    A = _ad()
    groups = NAMES + ['Zernike Coefficients (Fourier (Chebyshev ()))', 'Chebyshev-Fourier Coefficients (raw) ()',
                      'Chebyshev-Fourier Coefficients (raw) (Fourier ())', 'Comb Moments (raw) (Chebyshev ())',
                      'Gabor Textures (raw) ()']
    plan = A.DifferentiablePlan(groups)
    ref_groups = [A._split_raw(g) for g in groups]
    x = batch([9, 10], 48, 56)
    xg = x.clone().requires_grad_(True)
    f = plan(xg)
    assert f.shape == (2, plan.n_features) and plan.n_features == 4 * 72 + 400 + 2 * 529 + 320 + 8
    rng = np.random.default_rng(4)
    c = torch.from_numpy(rng.standard_normal(plan.n_features)).to(DEV)
    (f * c).sum().backward()
    for i in range(2):
        xr = x[i].cpu().clone().requires_grad_(True)
        fr = R.features(xr, ref_groups)
        assert rel(f[i].detach().cpu().numpy(), fr.detach().numpy()) < 1e-9
        assert rel(f[i].detach().cpu().numpy(), orc_features(x[i].cpu().numpy(), ref_groups)) < 1e-9
        (fr * c.cpu()).sum().backward()
        assert rel(xg.grad[i].cpu().numpy(), xr.grad.numpy()) < 1e-9


def test_raw_groups_reject_small_planes():
    import wndcharm
    A = _ad()
    plan = A.DifferentiablePlan(['Chebyshev-Fourier Coefficients (raw) ()'])
    with pytest.raises(wndcharm.WndGpuError):
        plan(batch([1], 20, 20))  # NN = 2 * min(11, (20 - 1) // 2) + 1 = 19: 361 magnitudes, not 529


def test_masked_groups_are_exact_and_carry_no_gradient():
    import wndcharm
    A = _ad()
    groups = ['Zernike Coefficients ()', 'Chebyshev Coefficients (Fourier ())', 'Haralick Textures ()']
    with pytest.raises(wndcharm.WndGpuError):
        A.DifferentiablePlan(groups)
    plan = A.DifferentiablePlan(groups, masked=True)
    x = batch([12, 13], 64, 64)
    xg = x.clone().requires_grad_(True)
    f = plan(xg)
    want = wndcharm.compute_features([x[i].cpu().numpy() for i in range(2)],
                                     _plan_of(wndcharm, groups[1:]))
    assert np.array_equal(f[:, 72:].detach().cpu().numpy(), want)
    m = plan.differentiable_mask()
    assert m[:72].all() and not m[72:].any()
    f[:, 72:].sum().backward()
    assert xg.grad is not None and (xg.grad == 0).all()


def _plan_of(wndcharm, groups):
    p = wndcharm.FeatureComputationPlan('masked')
    for g in groups:
        p.add(g)
    p.finalize()
    return p


def _model(shape, groups, n_classes=3, per_class=5, seed=300):
    A = _ad()
    plan = A.DifferentiablePlan(groups)
    imgs, cls = [], []
    for c in range(n_classes):
        from wndcharm import synthetic
        for k in range(per_class):
            imgs.append(synthetic.image(seed + 100 * c + k, [6, 12, 24, 48][c % 4], shape).astype(np.float64) / 4095)
            cls.append(c + 1)
    feats = plan(torch.from_numpy(np.stack(imgs)).to(DEV)).cpu().numpy()
    return plan, A.WNDModel(feats, cls, n_classes, used_fraction=0.5), feats, np.asarray(cls, np.int32)


def test_wnd_probabilities_and_gradient_match_reference():
    shape = (40, 48)
    plan, model, feats, cls = _model(shape, NAMES)
    x = batch([41, 42, 43], *shape)
    xg = x.clone().requires_grad_(True)
    p = model.probabilities(plan(xg))
    pn = p.detach().cpu().numpy()
    t = torch.from_numpy
    mins, maxs = model.mins.cpu().numpy(), model.maxs.cpu().numpy()
    data, w = model.train.cpu().numpy(), model.weights.cpu().numpy()
    fx = plan(x).cpu().numpy()
    for i in range(3):
        assert rel(pn[i], orc_probabilities(fx[i], data, cls, w, mins, maxs, 3)) < 1e-9
    k = int(np.argmin(pn[0]))
    p[:, k].sum().backward()
    ref_groups = [_ad()._split_raw(g) for g in NAMES]
    for i in range(3):
        xr = x[i].cpu().clone().requires_grad_(True)
        pr = R.wnd5(R.features(xr, ref_groups)[None], t(data), t(cls), t(w), t(mins), t(maxs), 3)[0]
        pr[k].backward()
        assert rel(xg.grad[i].cpu().numpy(), xr.grad.numpy()) < 1e-7


def _fd_deviation(plan, model, x, k, g, pix, h, order=2):
    """central differences minus g at pix, and g there:
    order 2: (f(+h) - f(-h)) / 2h; order 4: (-f(+2h) + 8f(+h) - 8f(-h) + f(-2h)) / 12h"""
    shape = x.shape[1:]
    steps = [h, -h] if order == 2 else [2 * h, h, -h, -2 * h]
    coef = [1, -1] if order == 2 else [-1, 8, -8, 1]
    den = 2 * h if order == 2 else 12 * h
    ns = len(steps)
    pert = x.repeat(ns * len(pix), 1, 1)
    ys, xs = np.unravel_index(pix, shape)
    idx = torch.arange(len(pix), device=DEV)
    yt, xt = torch.from_numpy(ys).to(DEV), torch.from_numpy(xs).to(DEV)
    for j, st in enumerate(steps):
        pert[ns * idx + j, yt, xt] += st
    with torch.no_grad():
        pp = torch.cat([model.probabilities(plan(pert[i:i + 64])) for i in range(0, len(pert), 64)])[:, k]
    pp = pp.cpu().numpy().reshape(len(pix), ns)
    fd = (pp @ np.asarray(coef, np.float64)) / den
    gs = g[ys, xs]
    return fd - gs, gs


def _fd_errors(*args, **kw):
    """max- and norm-relative deviation of central differences from g at pix"""
    dev, gs = _fd_deviation(*args, **kw)
    return np.max(np.abs(dev)) / np.max(np.abs(gs)), np.linalg.norm(dev) / np.linalg.norm(gs)


def _near_zero_frequencies(x, bound):
    """one (u, v) per conjugate pair of frequencies of the plane with |F(u, v)| < bound"""
    F = np.fft.fft2(x)
    H, W = x.shape
    out = []
    for u, v in zip(*np.nonzero(np.abs(F) < bound)):
        if ((-u) % H, (-v) % W) not in out:
            out.append((int(u), int(v)))
    return F, out


def _kink_terms(F, freqs, pix, h, order=2):
    """the stencil's error on |F(u, v)| alone at each pixel, one column per frequency pair: a step s at
    pixel p moves F(u, v) by s * e_p, e_p = exp(-2 pi i (u y / H + v x / W)); the stencil applied to
    |F + s e_p| minus the derivative Re(conj(F) e_p) / |F|.  The partner (-u, -v) of a real plane has
    the conjugate F and e_p, so the same column."""
    H, W = F.shape
    ys, xs = np.unravel_index(pix, (H, W))
    steps = [h, -h] if order == 2 else [2 * h, h, -h, -2 * h]
    coef = [1, -1] if order == 2 else [-1, 8, -8, 1]
    den = 2 * h if order == 2 else 12 * h
    cols = []
    for u, v in freqs:
        f0 = F[u, v]
        e = np.exp(-2j * np.pi * (u * ys / H + v * xs / W))
        d = np.real(np.conj(f0) * e) / abs(f0)
        cols.append(sum(c * np.abs(f0 + s * e) for c, s in zip(coef, steps)) / den - d)
    return np.stack(cols, axis=1)


def _fd_case(groups, seed_model, seed_img, pix_seed):
    shape = (512, 512)
    plan, model, _, _ = _model(shape, groups, n_classes=3, per_class=4, seed=seed_model)
    x = batch([seed_img], *shape)
    xg = x.clone().requires_grad_(True)
    p = model.probabilities(plan(xg))[0]
    k = int(torch.argmin(p))
    p[k].backward()
    g = xg.grad[0].cpu().numpy()
    pix = np.random.default_rng(pix_seed).choice(shape[0] * shape[1], 256, replace=False)
    return plan, model, x, k, g, pix


SMOOTH = ['Zernike Coefficients ()', 'Zernike Coefficients (Chebyshev ())', 'Chebyshev Coefficients (raw) ()',
          'Chebyshev-Fourier Coefficients (raw) ()', 'Comb Moments (raw) ()', 'Gabor Textures (raw) ()']


def test_gradient_matches_finite_differences_512():
    """pixels -> features -> WND5 probability of a class on 512 x 512 planes:
    central differences of the GPU forward at 256 random pixels.

    Without the Fourier node every feature is smooth and the two-point
    difference at h = 1e-3 agrees to 1e-4.  |F| is not differentiable where a
    frequency's value is 0, and the Fourier case's plane has a conjugate pair
    with |F| = 9.4e-4 < h: a 1e-3 step can carry F(u, v) through 0.  The
    deviation at h = 1e-3 is therefore modelled as the stencils' exact error
    on |F| at the frequencies with |F| < 10h (one weight per pair, least
    squares over the 256 pixels).  With those terms removed the rest must be
    under 2e-5, and the same weights must predict, with no new parameter, the
    deviations at h = 1e-4 and of the fourth-order stencil at h = 1e-3 (a
    smooth truncation error would fall 100x and ~1e4x; the kink gives ~98x
    and ~1.8x).  The gradient itself must agree at h = 1e-4."""
    args = _fd_case(SMOOTH, 900, 78, 13)
    s3, t3 = _fd_errors(*args, 1e-3)
    args = _fd_case(NAMES, 800, 77, 12)
    dev3, gs = _fd_deviation(*args, 1e-3)
    dev4, _ = _fd_deviation(*args, 1e-4)
    devq, _ = _fd_deviation(*args, 1e-3, order=4)
    F, freqs = _near_zero_frequencies(args[2][0].cpu().numpy(), 1e-2)
    pix = args[5]
    w = np.linalg.lstsq(_kink_terms(F, freqs, pix, 1e-3), dev3, rcond=None)[0]
    rest = dev3 - _kink_terms(F, freqs, pix, 1e-3) @ w
    pred4 = _kink_terms(F, freqs, pix, 1e-4) @ w
    predq = _kink_terms(F, freqs, pix, 1e-3, order=4) @ w
    gmax, gnorm = np.max(np.abs(gs)), np.linalg.norm(gs)
    mx = lambda d: np.max(np.abs(d)) / gmax
    nr = lambda d: np.linalg.norm(d) / gnorm
    print('finite differences at 256 pixels (max / norm relative): smooth families h=1e-3 %.3e / %.3e; '
          'with Fourier: 2-point h=1e-3 %.3e / %.3e, h=1e-4 %.3e / %.3e, 4th-order h=1e-3 %.3e / %.3e; '
          'frequencies with |F| < 1e-2: %s (|F| %s); h=1e-3 without their terms %.3e / %.3e; '
          'predicted h=1e-4 %.3e / %.3e (miss %.3e), 4th-order %.3e / %.3e (miss %.3e)'
          % (s3, t3, mx(dev3), nr(dev3), mx(dev4), nr(dev4), mx(devq), nr(devq), freqs,
             ['%.3e' % abs(F[u, v]) for u, v in freqs], mx(rest), nr(rest), mx(pred4), nr(pred4),
             nr(dev4 - pred4) / nr(dev4), mx(predq), nr(predq), nr(devq - predq) / nr(devq)))
    assert s3 <= 1e-4 and t3 <= 1e-4
    assert mx(dev4) <= 1e-5 and nr(dev4) <= 1e-5
    assert freqs and mx(rest) <= 2e-5 and nr(rest) <= 2e-5
    assert np.linalg.norm(dev4 - pred4) <= 0.1 * np.linalg.norm(dev4)
    assert np.linalg.norm(devq - predq) <= 0.1 * np.linalg.norm(devq)
