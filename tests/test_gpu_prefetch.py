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


def test_tiny_table_budget_evicts_and_stays_exact():
    """a context whose size-dependent table cache holds almost nothing
    (WNDGPU_TABLE_CACHE_MB=1): every batch evicts the previous batches'
    tables (stream-ordered frees) while prefetched ones are pinned for the
    batch they were queued for; features equal a default context's"""
    import os
    shapes = [[(97, 181), (120, 64)], [(333, 317), (64, 250)], [(97, 181), (301, 300)], [(80, 90)]]
    batches = [[synthetic.image(3200 + 10 * b + k, 5, shp) for k, shp in enumerate(bs)] for b, bs in enumerate(shapes)]
    plan = wndcharm.StdFeatureComputationPlans.getFeatureSetLong()
    want = _fresh_context_features(batches, plan)
    lib = wndcharm._load()
    old = wndcharm._context()
    os.environ['WNDGPU_TABLE_CACHE_MB'] = '1'
    c = ctypes.c_void_p()
    try:
        assert lib.wndgpu_create(0, ctypes.byref(c)) == 0
    finally:
        os.environ.pop('WNDGPU_TABLE_CACHE_MB')
    wndcharm._ctx = c
    try:
        print('tiny table budget', [[im.shape for im in b] for b in batches], flush=True)
        wndcharm.prefetch_tables(batches[0], plan)
        got = []
        for k, b in enumerate(batches):
            if k + 1 < len(batches):
                wndcharm.prefetch_tables(batches[k + 1], plan)
            got.append(wndcharm.compute_features(b, plan))
        wndcharm.prefetch_tables(batches[0], plan)  # queued, never taken: dropped at teardown
    finally:
        wndcharm._ctx = old
        lib.wndgpu_destroy(c)
    for g, w in zip(got, want):
        assert np.array_equal(g, w, equal_nan=True)
