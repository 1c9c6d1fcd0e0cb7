"""Batch graphs (wndgpu_set_graph_mode): a repeated batch signature is captured
once into a hipGraph and replayed.  Every replay must equal the eager
computation of the same images bit for bit, with new pixel data each time
(the image and output addresses reach the kernels only through the staged
descriptors the graph copies), with eager batches of other signatures run
between the replays, and under the schedule autotuner."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module', autouse=True)
def torch_first():
    """torch's HIP runtime is initialised before the library's context, as in
    the other GPU test modules"""
    import torch
    torch.zeros(1, device='cuda')
    yield


def _images(seed, shapes):
    from wndcharm import synthetic
    return [synthetic.image(seed + k, 8, s) for k, s in enumerate(shapes)]


SHAPES_A = [(64, 64)] * 3 + [(48, 80), (80, 48)]
SHAPES_B = [(72, 56), (56, 72)]


@pytest.fixture
def graph_mode():
    import wndcharm
    yield wndcharm.set_graph_mode
    wndcharm.set_graph_mode(0)


@pytest.mark.parametrize('mode', [1, 2])
def test_replays_equal_eager(graph_mode, mode):
    import wndcharm
    plan = wndcharm.StdFeatureComputationPlans.getFeatureSetLong()
    batches = [_images(100 * k, SHAPES_A) for k in range(5)] + [_images(900, SHAPES_B)]
    graph_mode(0)
    eager = [wndcharm.compute_features(b, plan) for b in batches]
    graph_mode(mode)
    c0, r0, _, _ = wndcharm.graph_stats()
    order = [0, 1, 2, 5, 3, 4, 2, 5, 0]  # signature A, A (captured), A (replay), B, ...
    for k in order:
        got = wndcharm.compute_features(batches[k], plan)
        assert np.array_equal(got, eager[k], equal_nan=True), 'batch %d differs from its eager result' % k
    c1, r1, cached, sched = wndcharm.graph_stats()
    assert c1 - c0 == (2 if mode == 2 else 1) * 2  # A and B, each once per schedule
    assert r1 - r0 == 7  # every call from a signature's second on launches its graph: A 6, B 1
    assert cached >= 2
    if mode == 2:
        assert sched in (0, 1)


def test_device_io_replays(graph_mode):
    import ctypes
    import torch
    import wndcharm
    plan = wndcharm.StdFeatureComputationPlans.getFeatureSetLong()
    lib, ctx = wndcharm._load(), wndcharm._context()
    batches = [_images(300 + 10 * k, SHAPES_A) for k in range(4)]
    graph_mode(0)
    eager = [wndcharm.compute_features(b, plan) for b in batches]
    graph_mode(1)
    for k in [0, 1, 2, 3, 1]:
        dev = [torch.from_numpy(im.view(np.int16)).cuda() for im in batches[k]]
        out = torch.zeros((len(dev), plan.n_features), dtype=torch.float64, device='cuda')
        st = (wndcharm._Image * len(dev))()
        for i, t in enumerate(dev):
            st[i] = wndcharm._Image(ctypes.c_void_p(t.data_ptr()), t.shape[0], t.shape[1], t.shape[1], 0)
        torch.cuda.synchronize()
        assert lib.wndgpu_compute_device(ctx, plan.handle, st, len(dev), ctypes.c_void_p(out.data_ptr()), None) == 0
        assert np.array_equal(out.cpu().numpy(), eager[k], equal_nan=True)
