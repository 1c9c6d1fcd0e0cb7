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
* cf        ChebyshevFourierCoefficients before its histogram (Downsample above
            300 x 300, src/cmatrix.cpp:579-720 and 1227-1240; ChebyshevFourier2D,
            src/transforms/ChebyshevFourier.cpp:59-160): the NN x NN magnitudes.
* comb      CombFirstFourMoments before its histograms
            (src/statistics/CombFirst4Moments.cpp:78-217): the mean, std, skew and
            kurtosis of each of 4 x 20 pixel lines.
* gabor     the 8 energy maps GaborTextureFilters2D thresholds (GaborEnergy,
            src/textures/gabor.cpp:166-231); "Gabor Textures (raw)" is their means.
* wnd5      signatures::normalize (src/wndchrm_src/signatures.cpp:196-215) and
            TrainingSet::classify2 (src/wndchrm_src/TrainingSet.cpp:2429-2586).
"""
import functools
import math

import numpy as np
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


def downsample_matrix(L, ratio):
    """dense weights [int(ratio L)][L] of one pass of ImageMatrix::Downsample
    (src/cmatrix.cpp:579-720): output k averages [x_k, x_k + d) with
    fractional end weights, x_k by the sequential x += d, d = 1 / ratio"""
    r = min(ratio, 1.0)
    d = 1 / r
    rows = []
    x = 0.0
    while x < L:
        row = np.zeros(L)
        a = int(math.floor(x))
        frac = math.ceil(x) - x
        if frac > 0 and a < L:
            row[a] += frac
        a = int(math.ceil(x))
        while a < math.floor(x + d):
            if a < L:
                row[a] += 1
            a += 1
        frac = x + d - math.floor(x + d)
        if frac > 0 and a < L:
            row[a] += frac
        rows.append(row / d)
        x += d
    return torch.from_numpy(np.stack(rows[:int(r * L)]))


@functools.lru_cache(maxsize=8)
def cf_geometry(h, w):
    """in-disk pixels (row-major indices, x-outer order) and the basis tables
    of ChebyshevFourier2D (src/transforms/ChebyshevFourier.cpp:59-127)"""
    m, n = h, w
    kk, tn, fc, fs = [], [], [], []
    N = min(11, (min(m, n) - 1) // 2)
    NN = 2 * N + 1
    for x in range(n):
        xi = -1.0 + x * (2.0 / (n - 1))
        for y in range(m):
            yi = -1.0 + y * (2.0 / (m - 1))
            r = math.sqrt(xi * xi + yi * yi)
            f = -1 * math.atan2(yi, xi)
            if not r < 1:
                continue
            kk.append(y * n + x)
            xx = min(max(r * 2 - 1, -1.0), 1.0)
            t = [math.cos(a * math.acos(xx)) for a in range(NN)]
            t[0] = 0.5
            tn.append(t)
            mf = np.arange(NN) - N
            ft = np.where(mf == 0, 0.5, 1.0)
            fc.append(ft * np.cos(mf * f))
            fs.append(ft * np.sin(-1 * mf * f))
    t = lambda a: torch.from_numpy(np.asarray(a, np.float64))
    return torch.tensor(kk), t(tn), t(fc), t(fs), NN


def cf_magnitudes(img):
    """the NN x NN magnitudes ChebyshevFourierCoefficients histograms
    (src/FeatureAlgorithms.cpp:54-69, src/cmatrix.cpp:1227-1240)"""
    h, w = img.shape
    if w * h > 300 * 300:
        rr = min(300.0 / w, 300.0 / h)
        img = downsample_matrix(h, rr) @ img @ downsample_matrix(w, rr).t()
        h, w = img.shape
    kk, tn, fc, fs, NN = cf_geometry(h, w)
    v = img.reshape(-1)[kk]
    cr = (fc * v[:, None]).t() @ tn
    ci = (fs * v[:, None]).t() @ tn
    return torch.sqrt(cr * cr + ci * ci).reshape(-1)


def comb_lines(h, w):
    """pixel indices of the 4 x 20 lines of CombFirst4Moments2D
    (src/statistics/CombFirst4Moments.cpp:78-217): positions ii = round(mp),
    mp = 1 - m, 1 - m + m/10, ... while mp <= m (m = height, step >= 1);
    comb 0: |I + ii - J| < 1, comb 1: |I + ii - (n - x)| < 1, comb 2:
    |J + ii - n/2| < 1, comb 3: |I + ii - m/2| < 1 (I = y+1, J = x+1, integer halves)"""
    m, n = h, w
    n2, m2 = n // 2, m // 2
    step = max(m / 10.0, 1.0)
    yy, xx = np.mgrid[0:m, 0:n]
    I, J, J1 = yy + 1.0, xx + 1.0, n - xx
    lines = []
    for comb in range(4):
        mp, per = 1.0 - m, []
        while mp <= m:
            ii = float(math.floor(abs(mp) + 0.5) * (1 if mp >= 0 else -1))  # C round: half away from zero
            if comb == 0:
                sel = np.abs(I + ii - J) < 1
            elif comb == 1:
                sel = np.abs(I + ii - J1) < 1
            elif comb == 2:
                sel = np.abs(J + ii - n2) < 1
            else:
                sel = np.abs(I + ii - m2) < 1
            if len(per) < 20:
                per.append(torch.from_numpy(np.flatnonzero(sel.reshape(-1))))
            mp += step
        while len(per) < 20:
            per.append(torch.zeros(0, dtype=torch.long))
        lines.append(per)
    return lines


def moments4(v):
    """mean, std, skew, kurt of Moments4 (src/statistics/Moments.h), two-pass"""
    n = v.numel()
    if n == 0:
        z = v.sum() * 0
        return [z, z, z, z]
    mu = v.mean()
    d = v - mu
    m2, m3, m4 = (d * d).sum(), (d ** 3).sum(), (d ** 4).sum()
    z = m2 * 0
    std = torch.sqrt(m2 / (n - 1)) if n > 2 else z
    skew = math.sqrt(n) * m3 / m2 ** 1.5 if n > 3 else z
    kurt = n * m4 / (m2 * m2) if n > 4 else z
    return [mu, std, skew, kurt]


def comb_raw(img):
    """the 4 x 4 x 20 line moments that CombFirstFourMoments histograms,
    [comb][mean, std, skew, kurt][position]"""
    flat = img.reshape(-1)
    out = []
    for per in comb_lines(*img.shape):
        mom = [moments4(flat[ix]) for ix in per]
        for a in range(4):
            out += [mom[b][a] for b in range(20)]
    return torch.stack(out)


def gabor_kernels():
    """the 8 complex 38 x 38 kernels of GaborTextureFilters2D (src/textures/gabor.cpp:103-162,
    201-231): gamma 0.5, sig2lam 0.56, theta = 3.14159265/2, f0 = 0.1 then 1..7, each
    divided by the sum of its magnitudes"""
    n, gamma, sig2lam, theta = 38, 0.5, 0.56, 3.14159265 / 2
    t = np.arange(n + 1, dtype=np.float64)[:n] - (n // 2)
    ty, tx = np.meshgrid(t, t, indexing='ij')
    out = []
    for f0 in (0.1, 1, 2, 3, 4, 5, 6, 7):
        lam = 2 * math.pi / f0
        sig = sig2lam * lam
        xte = tx * math.cos(theta) + ty * math.sin(theta)
        yte = ty * math.cos(theta) - tx * math.sin(theta)
        ge = np.exp(-1 * (xte * xte + gamma * gamma * yte * yte) / (2 * sig * sig))
        g = ge * np.cos(xte * f0) + 1j * ge * np.sin(xte * f0)
        out.append(g / np.abs(g).sum())
    return np.stack(out)


def gabor_energy(img):
    """the 8 energy maps of GaborEnergy (src/textures/gabor.cpp:166-195): the
    full convolution with each kernel, cropped at ceil(38/2) = 19 -> [8, h, w]"""
    h, w = img.shape
    G = gabor_kernels()
    # E(b, a) = |sum_{j,i} G[j][i] I[b + 19 - j][a + 19 - i]|: a correlation with the flipped kernel
    Gf = torch.from_numpy(np.ascontiguousarray(G[:, ::-1, ::-1]))
    x = torch.nn.functional.pad(img[None, None], (18, 19, 18, 19))
    re = torch.nn.functional.conv2d(x, Gf.real[:, None].contiguous())[0]
    im = torch.nn.functional.conv2d(x, Gf.imag[:, None].contiguous())[0]
    return torch.sqrt(re * re + im * im)


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
        elif alg == 'Chebyshev-Fourier Coefficients (raw)':
            out.append(cf_magnitudes(p))
        elif alg == 'Comb Moments (raw)':
            out.append(comb_raw(p))
        elif alg == 'Gabor Textures (raw)':
            out.append(gabor_energy(p).mean((1, 2)))
        else:
            raise ValueError(alg)
    return torch.cat(out)
