"""FP64 reference of the differentiable path, written with plain torch ops on
the CPU (test infrastructure, never the product path): its forward values are
pinned to the oracle in tests/test_ad_reference.py, its gradients (torch
autograd) to central finite differences of the oracle, and the GPU adjoint
kernels are then compared with these gradients.

Definitions follow the reference sources:
* fourier   ImageMatrix::fft2 (src/cmatrix.cpp:1107-1200): |DFT2| of the plane.
* chebyshev ImageMatrix::ChebyshevTransform -> Chebyshev2D
            (src/cmatrix.cpp:1203-1222, src/statistics/chebyshev.cpp:39-165):
            out[j][k] = 1/4 sum_x sum_y I[y][x] Tx[x][j] Ty[y][k], T[a][j] =
            cos(j acos(2(a+1)/L - 1)) * (1/L for j = 0, else 2/L), T[a][0] = 1/L.
* zernike   mb_zernike2D (src/textures/zernike/zernike.cpp:207-341): order 15,
            the angle of repetition m taken as (m+1) theta as the reference's
            COST / SINT tables hold it.
* wnd5      signatures::normalize (src/wndchrm_src/signatures.cpp:196-215) and
            TrainingSet::classify2 (src/wndchrm_src/TrainingSet.cpp:2429-2586).
"""
import math

import torch

F64 = torch.float64
EPS = 2.2204460492503131e-16


def fourier(img):
    return torch.fft.fft2(img).abs()


def cheb_basis(L, N):
    a = torch.arange(L, dtype=F64)
    x = 2 * (a + 1) / L - 1
    j = torch.arange(N, dtype=F64)
    T = torch.cos(j[None, :] * torch.acos(torch.clamp(x, -1, 1))[:, None])
    T[:, 0] = 1.0
    s = torch.full((N,), 2.0 / L, dtype=F64)
    s[0] = 1.0 / L
    return T * s[None, :]


def chebyshev(img, N=None):
    h, w = img.shape
    if N is None:
        N = min(h, w)
    bx, by = cheb_basis(w, N), cheb_basis(h, N)
    return 0.25 * (bx.t() @ img.t() @ by)


def _radial_coeffs(n, m):
    K = (n - m) // 2
    return [((-1) ** s) * math.factorial(n - s) /
            (math.factorial(s) * math.factorial((n + m) // 2 - s) * math.factorial(K - s)) for s in range(K + 1)]


def zernike(img, L=15):
    h, w = img.shape
    S = img.sum()
    ii = torch.arange(1, w + 1, dtype=F64)
    jj = torch.arange(1, h + 1, dtype=F64)
    xc = (img * ii[None, :]).sum() / S
    yc = (img * jj[:, None]).sum() / S
    rad = float(min(w, h))
    X = ((ii[None, :] - xc) / rad).expand(h, w)
    Y = ((jj[:, None] - yc) / rad).expand(h, w)
    r = torch.sqrt(X * X + Y * Y)
    mask = ((r >= EPS) & (r <= 1.0)).detach()
    Xs = torch.where(mask, X, torch.ones_like(X))
    Ys = torch.where(mask, Y, torch.zeros_like(Y))
    rs = torch.where(mask, r, torch.ones_like(r))
    theta = torch.atan2(Ys, Xs)
    f = torch.where(mask, img / S, torch.zeros_like(img))
    out = []
    for n in range(L + 1):
        for m in range(n + 1):
            if (n - m) % 2:
                continue
            R = sum(c * rs ** (n - 2 * s) for s, c in enumerate(_radial_coeffs(n, m)))
            k = (n + 1) / math.pi * f * R
            re = (k * torch.cos((m + 1) * theta)).sum()
            im = -(k * torch.sin((m + 1) * theta)).sum()
            out.append(torch.sqrt(re * re + im * im))
    return torch.stack(out)


def normalize(x, mins, maxs):
    rng = maxs - mins
    zero = torch.isnan(x) | (x < mins) | (rng < EPS)
    lin = 100 * (x - mins) / torch.where(rng < EPS, torch.ones_like(rng), rng)
    v = torch.where(x > maxs, torch.full_like(x, 100.0), lin)
    return torch.where(zero, torch.zeros_like(x), v)


def wnd5(x, train, classes, weights, mins, maxs, n_classes):
    """probabilities [n, C] of raw vectors x [n, F]"""
    v = normalize(x, mins, maxs)
    d = v[:, None, :] - train[None, :, :]
    keep = ((d.abs() >= EPS) & (weights[None, None, :] >= EPS)).detach()
    dist = torch.where(keep, (weights ** 2)[None, None, :] * d * d, torch.zeros_like(d)).sum(-1)
    ok = (dist >= EPS).detach()
    s = torch.where(ok, torch.where(ok, dist, torch.ones_like(dist)) ** -5, torch.zeros_like(dist))
    cols = []
    for c in range(1, n_classes + 1):
        sel = (classes == c)
        cnt = ok[:, sel].sum(1)
        tot = s[:, sel].sum(1)
        cols.append(torch.where(cnt > 0, tot / torch.clamp(cnt, min=1), torch.zeros_like(tot)))
    sb = torch.stack(cols, 1)
    return sb / sb.sum(1, keepdim=True)


TRANSFORMS = {'Fourier': fourier, 'Chebyshev': chebyshev}


def node(img, chain):
    for t in chain:
        img = TRANSFORMS[t](img)
    return img


def features(img, groups):
    """feature vector of one plane for (algorithm, chain) groups"""
    out = []
    for alg, chain in groups:
        p = node(img, chain)
        if alg == 'Zernike Coefficients':
            out.append(zernike(p))
        elif alg == 'Chebyshev Coefficients (raw)':
            out.append(chebyshev(p, min(20, *p.shape)).reshape(-1))
        else:
            raise ValueError(alg)
    return torch.cat(out)
