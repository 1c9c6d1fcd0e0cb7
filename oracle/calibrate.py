"""Speed calibration of the oracle (the CPU restatement timed as the
benchmark's cpu_baseline) against the reference's own code, for the units of
the feature path whose reference sources build here (oracle/ref ->
oracle/_ref/libwndref.so): Haralick textures with f14 (CVIPtexture.cpp +
Eigen 3.1.2), Radon projections (radon.cpp), the Symlet-5 wavelet
(Wavelet.cpp, convolution.cpp) and the Chebyshev transform (chebyshev.cpp).
The whole reference feature path needs FFTW3 and generated headers that this
image lacks, so no whole-vector ratio exists.  Test infrastructure: run in
the build container, single-threaded, on the 256x256 configuration images.

    python oracle/calibrate.py [n_images]

Prints one JSON line: per unit the oracle's and the reference's seconds per
image, their ratio, and the unit's share of the oracle's long-set time."""
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, '..', 'wnd-charm_amd'))
P, I = ctypes.c_void_p, ctypes.c_int


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def main():
    from wndcharm import synthetic
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    orc = ctypes.CDLL(os.path.join(HERE, 'liboracle.so'))
    ref = ctypes.CDLL(os.path.join(HERE, '_ref', 'libwndref.so'))
    orc.orc_transform.argtypes = [I, P, I, I, P, P, P]
    orc.orc_algorithm.argtypes = [I, P, I, I, P]
    orc.orc_features.argtypes = [I, P, I, I, P]
    orc.orc_haralick_grays.argtypes = [P, I, I, P]
    orc.orc_radon_proj.argtypes = [P, I, I, P, P]
    ref.ref_texture.argtypes = [P, I, I, I, P]
    ref.ref_radon.argtypes = [P, I, I, I, I, P]
    ref.ref_wavelet.argtypes = [P, I, I, P, P, P]
    ref.ref_cheb1d.argtypes = [P, I, I, I, P]
    imgs = [im.astype(np.float64) for im in synthetic.config_images('long_256', limit=n)[0]]
    h, w = imgs[0].shape
    t = {}

    def clock(name, f, planes=None):
        """best of three passes, seconds per image"""
        best = 1e30
        for _ in range(3):
            t0 = time.perf_counter()
            for im in (planes if planes is not None else imgs):
                f(im)
            best = min(best, time.perf_counter() - t0)
        t[name] = best / len(imgs)

    def transform(kind, im):
        hh, ww = im.shape
        o = np.zeros((2 * ((hh + 9) // 2) + 2) * (2 * ((ww + 9) // 2) + 2))
        a, b = ctypes.c_int(), ctypes.c_int()
        orc.orc_transform(kind, ptr(np.ascontiguousarray(im)), hh, ww, ptr(o), ctypes.byref(a), ctypes.byref(b))
        return o[:a.value * b.value].reshape(a.value, b.value).copy()

    # the 12 nodes Haralick and Radon run on in the long set (transform chains,
    # innermost first): (), C, C(F), C(W), E, F, F(C), F(E), F(W), W, W(E), W(F)
    chains = [[], [1], [0, 1], [2, 1], [3], [0], [1, 0], [3, 0], [2, 0], [2], [3, 2], [0, 2]]
    nodes = []
    for im in imgs:
        for ch in chains:
            p = im
            for k in ch:
                p = transform(k, p)
            nodes.append(np.ascontiguousarray(p))

    HARALICK, RADON, T_CHEB, T_WAV = 3, 7, 1, 2
    big = np.zeros((2 * ((h + 9) // 2) + 2) * (2 * ((w + 9) // 2) + 2))
    oh, ow = ctypes.c_int(), ctypes.c_int()
    out28 = np.zeros(64)
    clock('haralick_oracle', lambda im: orc.orc_algorithm(HARALICK, ptr(im), im.shape[0], im.shape[1], ptr(out28)),
          nodes)

    def har_ref(im):
        hh, ww = im.shape
        g = np.zeros(hh * ww, np.uint8)
        orc.orc_haralick_grays(ptr(im), hh, ww, ptr(g))  # the gray levels (exact in both)
        o = np.zeros(14)
        for ang in (0, 45, 90, 135):
            ref.ref_texture(ptr(g), hh, ww, ang, ptr(o))
    clock('haralick_reference', har_ref, nodes)
    proj = np.zeros(4 * (2 * (1100 + 1100) + 8))
    rs = ctypes.c_int()
    clock('radon_oracle', lambda im: orc.orc_radon_proj(ptr(im), im.shape[0], im.shape[1], ptr(proj),
                                                        ctypes.byref(rs)), nodes)

    def radon_ref(im):
        hh, ww = im.shape
        orc.orc_radon_proj(ptr(im), hh, ww, ptr(proj), ctypes.byref(rs))  # its bin count only
        r = rs.value
        cm = np.ascontiguousarray(im.T)
        t0 = time.perf_counter()
        ref.ref_radon(ptr(cm), hh, ww, -((r - 1) // 2), r, ptr(proj))
        radon_ref.own += time.perf_counter() - t0
    radon_ref.own = 0.0
    clock('radon_reference', radon_ref, nodes)
    t['radon_reference'] = radon_ref.own / 3 / len(imgs)
    clock('wavelet_oracle', lambda im: orc.orc_transform(T_WAV, ptr(im), h, w, ptr(big), ctypes.byref(oh),
                                                         ctypes.byref(ow)))
    wout = np.zeros((h + 20) * (w + 20))
    clock('wavelet_reference', lambda im: ref.ref_wavelet(ptr(im), h, w, ptr(wout), ctypes.byref(oh),
                                                          ctypes.byref(ow)))
    clock('chebyshev_oracle', lambda im: orc.orc_transform(T_CHEB, ptr(im), h, w, ptr(big), ctypes.byref(oh),
                                                           ctypes.byref(ow)))
    nn = min(h, w)
    a = np.zeros(h * nn)
    b = np.zeros(nn * nn)

    def cheb_ref(im):
        ref.ref_cheb1d(ptr(im), h, w, nn, ptr(a))
        tt = np.ascontiguousarray(a.reshape(h, nn).T)
        ref.ref_cheb1d(ptr(tt), nn, h, nn, ptr(b))
    clock('chebyshev_reference', cheb_ref)
    feats = np.zeros(2919)
    clock('long_set_oracle', lambda im: orc.orc_features(1, ptr(im), h, w, ptr(feats)))
    units = {}
    for u, mult in (('haralick', 1), ('radon', 1), ('wavelet', 5), ('chebyshev', 4)):
        o, rf = t[u + '_oracle'], t[u + '_reference']
        units[u] = {'oracle_s': round(o, 4), 'reference_s': round(rf, 4), 'oracle_over_reference': round(o / rf, 3),
                    'long_set_share': round(mult * o / t['long_set_oracle'], 3)}
    print(json.dumps({'images': '%d long-set configuration images %dx%d, one thread, best of 3' % (n, h, w),
                      'long_set_oracle_s_per_image': round(t['long_set_oracle'], 3), 'units': units,
                      'note': 'Haralick and Radon timed on the 12 node planes they run on per image; the wavelet '
                              'and Chebyshev transforms on the image, their share counting the 5 and 4 times the '
                              'long set computes them (approximate: their inputs differ in size)'}))


if __name__ == '__main__':
    main()
