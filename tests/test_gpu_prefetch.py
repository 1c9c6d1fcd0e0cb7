"""Size-dependent tables queued ahead of the compute call that needs them
(wndgpu_prefetch_tables): features equal, bit for bit, those of a context
that builds every table inside the call; the call times are reported."""
import ctypes

import numpy as np
import pytest

import wndcharm
from wndcharm import synthetic

pytestmark = pytest.mark.gpu


def _fresh_context_features(batches, plan):
    lib = wndcharm._load()
    old = wndcharm._context()
    c = ctypes.c_void_p()
    assert lib.wndgpu_create(0, ctypes.byref(c)) == 0
    wndcharm._ctx = c
    try:
        return [wndcharm.compute_features(b, plan) for b in batches]
    finally:
        wndcharm._ctx = old
        lib.wndgpu_destroy(c)


def test_prefetched_tables_same_features():
    # new geometries: Chebyshev bases of new sides, Chebyshev-Fourier tables
    # of new (downsampled above 300 x 300) plane sizes
    shapes = [(333, 317), (350, 290), (97, 181), (401, 260), (211, 77)]
    imgs = [synthetic.image(3001 + k, 7, shp) for k, shp in enumerate(shapes)]
    a, b = imgs[:2], imgs[2:]
    plan = wndcharm.StdFeatureComputationPlans.getFeatureSetLong()
    print('prefetch', [im.shape for im in a], [im.shape for im in b], flush=True)
    wndcharm.prefetch_tables(a, plan)
    wndcharm.prefetch_tables(b, plan)
    fa = wndcharm.compute_features(a, plan)
    host, table, dev = wndcharm.last_call_times()
    assert host >= table >= 0 and dev >= 0
    fb = wndcharm.compute_features(b, plan)
    ra, rb = _fresh_context_features([a, b], plan)
    assert np.array_equal(fa, ra, equal_nan=True)
    assert np.array_equal(fb, rb, equal_nan=True)


def test_prefetch_rejects_nothing_and_reads_no_pixels():
    """images that a compute would reject are skipped, not errors"""
    plan = wndcharm.StdFeatureComputationPlans.getFeatureSet()
    tiny = np.zeros((8, 8), np.uint16)
    ok = synthetic.image(3101, 5, (120, 90))
    wndcharm.prefetch_tables([tiny, ok], plan)
    got = wndcharm.compute_features([ok], plan)
    assert np.isfinite(got).any()
