"""Pixel loading.  uint16 planes with contiguous, 8-byte-aligned rows are
converted four pixels per thread; row strides wider than the image, sources
that are only 2-byte aligned and the other sample types take the per-pixel
path.  Every path must give the root plane of the float64 input exactly, so
the features are bit-identical to those of the float64 images.  The pixel
counts of the shapes cover every remainder mod 4 (the fast path's tail)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(40, 48), (33, 29), (31, 34), (35, 33)]  # 1920, 957, 1054, 1155 pixels: 0, 1, 2, 3 mod 4


@pytest.fixture(scope='module', autouse=True)
def torch_first():
    """torch's HIP runtime is initialised before the library's context"""
    import torch
    torch.zeros(1, device='cuda')
    yield


def _images():
    from wndcharm import synthetic
    return [synthetic.image(4100 + k, 8, s) for k, s in enumerate(SHAPES)]


def _reference(images, plan):
    import wndcharm
    return wndcharm.compute_features([im.astype(np.float64) for im in images], plan)


def test_host_uint16_equals_float64():
    import wndcharm
    plan = wndcharm.StdFeatureComputationPlans.getFeatureSetLong()
    images = _images()
    assert all(im.dtype == np.uint16 for im in images)
    print('host uint16 vs float64', SHAPES, flush=True)
    got = wndcharm.compute_features(images, plan)
    assert np.array_equal(got, _reference(images, plan), equal_nan=True)


@pytest.mark.parametrize('offset,pad', [(0, 0), (1, 0), (0, 3), (3, 5)],
                         ids=['aligned', 'offset2B', 'stride', 'offset6B_stride'])
def test_device_uint16_layouts(offset, pad):
    """device sources at a byte offset of 2 * offset into a torch buffer, rows
    pad pixels wider than the image"""
    import torch
    import wndcharm
    plan = wndcharm.StdFeatureComputationPlans.getFeatureSetLong()
    lib, ctx = wndcharm._load(), wndcharm._context()
    images = _images()
    print('device uint16 layout offset', offset, 'pad', pad, flush=True)
    bufs = []
    st = (wndcharm._Image * len(images))()
    for i, im in enumerate(images):
        h, w = im.shape
        stride = w + pad
        host = np.full(offset + h * stride, 0x5a5a, np.uint16)  # padding holds a value no feature may see
        host[offset:].reshape(h, stride)[:, :w] = im
        buf = torch.from_numpy(host.view(np.int16)).cuda()
        bufs.append(buf)
        st[i] = wndcharm._Image(ctypes.c_void_p(buf.data_ptr() + 2 * offset), h, w, stride, 0)
    out = torch.zeros((len(images), plan.n_features), dtype=torch.float64, device='cuda')
    torch.cuda.synchronize()
    assert lib.wndgpu_compute_device(ctx, plan.handle, st, len(images), ctypes.c_void_p(out.data_ptr()), None) == 0
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), _reference(images, plan), equal_nan=True)
