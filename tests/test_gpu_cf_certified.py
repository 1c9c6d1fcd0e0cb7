"""Chebyshev-Fourier histograms from matrix-core sums
(ChebyshevFourier.cpp:59-160 in the reference): the 529 coefficient sums of
an entry are a GEMM on the FP64 matrix cores in a different order from the
reference's sequential sums, every histogram bin is certified against an
a-priori bound on that difference, and an entry with an open bin is
recomputed in the reference's order.  The counts must equal the oracle's
whether an entry is certified or recomputed; WNDGPU_CF_SLACK widens the
bound so that the recomputation runs for some or all entries."""
import ctypes
import os

import numpy as np
import pytest

import wndcharm
from wndcharm import synthetic
from helpers import ALG_NAMES, orc_algorithm, orc_transform

pytestmark = pytest.mark.gpu
CF = ALG_NAMES.index('Chebyshev-Fourier Coefficients')
# node name -> oracle transform chain (innermost first): Fourier 0, Chebyshev 1, Wavelet 2, Edge 3
NODES = {'()': [], '(Fourier ())': [0], '(Chebyshev ())': [1], '(Wavelet ())': [2], '(Edge ())': [3],
         '(Fourier (Edge ()))': [3, 0], '(Fourier (Wavelet ()))': [2, 0], '(Wavelet (Edge ()))': [3, 2]}


def _plan():
    plan = wndcharm.FeatureComputationPlan('cf')
    for node in NODES:
        plan.add('Chebyshev-Fourier Coefficients ' + node)
    plan.finalize()
    return plan


def _want(img):
    rows = []
    for chain in NODES.values():
        p = img.astype(np.float64)
        for k in chain:
            p = orc_transform(k, p)
        rows.append(orc_algorithm(CF, p))
    return np.concatenate(rows)


@pytest.fixture
def slack_context(request):
    """a context whose certification bound is scaled by request.param"""
    lib = wndcharm._load()
    old = wndcharm._context()
    os.environ['WNDGPU_CF_SLACK'] = repr(float(request.param))
    c = ctypes.c_void_p()
    try:
        rc = lib.wndgpu_create(0, ctypes.byref(c))
    finally:
        os.environ.pop('WNDGPU_CF_SLACK')
    assert rc == 0
    wndcharm._ctx = c
    try:
        yield float(request.param)
    finally:
        wndcharm._ctx = old
        lib.wndgpu_destroy(c)


def _images():
    imgs, _ = synthetic.config_images('long_256', limit=6)
    imgs = list(imgs)
    imgs.append(synthetic.image(2001, 8, (16, 16)))          # NN = 15
    imgs.append(synthetic.image(2002, 5, (37, 90)))          # NN = 37 -> 23, ragged
    imgs.append(np.full((64, 64), 77, np.uint16))            # constant plane
    imgs.append(synthetic.image(2003, 9, (301, 300)))        # downsampled above 300 x 300
    return imgs


def _check(imgs, plan):
    got = wndcharm.compute_features(imgs, plan)
    opened = wndcharm.last_cf_open()
    assert opened.shape == (len(imgs) * len(NODES),)
    for i, img in enumerate(imgs):
        want = _want(img)
        assert np.array_equal(got[i], want), (i, img.shape, int((got[i] != want).sum()))
    return opened


def test_certified_counts_equal_oracle():
    imgs = _images()
    print('certified Chebyshev-Fourier on', [im.shape for im in imgs], flush=True)
    opened = _check(imgs, _plan())
    print('entries recomputed in the reference order: %d of %d' % (opened.sum(), opened.size), flush=True)
    # a bin is open only if a magnitude lies within ~1e-10 relative of a bin
    # edge; on these planes the constant one is the only candidate
    assert opened.sum() <= 8


@pytest.mark.parametrize('slack_context', [1e15], indirect=True)
def test_every_entry_recomputed(slack_context):
    imgs = _images()
    print('Chebyshev-Fourier, every entry open (slack %g)' % slack_context, flush=True)
    opened = _check(imgs, _plan())
    assert opened.all()


@pytest.mark.parametrize('slack_context', [3e8], indirect=True)
def test_some_entries_recomputed(slack_context):
    imgs = _images()
    print('Chebyshev-Fourier, partly open (slack %g)' % slack_context, flush=True)
    opened = _check(imgs, _plan())
    print('entries recomputed: %d of %d' % (opened.sum(), opened.size), flush=True)
