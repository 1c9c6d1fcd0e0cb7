"""wndgpu_ad_wnd_forward_noalloc (WND5 probabilities of raw feature vectors
with its work in the library's static device array) against
wndgpu_ad_wnd_forward with a caller-owned work buffer: identical
probabilities on two calls in a row, the second with fewer test vectors and
a smaller training set, so stale scratch from the first call would show;
sizes above the array are refused before any GPU work."""
import ctypes

import numpy as np
import pytest

import wndcharm

pytestmark = pytest.mark.gpu

_P = ctypes.c_void_p
_I = ctypes.c_int


def _lib():
    from wndcharm import autograd
    lib = autograd._ad_lib()
    f = lib.wndgpu_ad_wnd_forward_noalloc
    f.restype = _I
    f.argtypes = [_P, _P, _P, _I, _I, _I, _P, _P, _P, _P, _I, _P, _P]
    return lib


def _space(seed, n_per_class, n_classes, n_features):
    rng = np.random.default_rng(seed)
    centres = rng.normal(0, 1, (n_classes, n_features)) * rng.uniform(0.5, 3, n_features)
    raw = np.concatenate([centres[c] + rng.normal(0, 1, (n_per_class, n_features)) for c in range(n_classes)])
    raw[:, :5] = np.round(raw[:, :5] * 10)  # a few integer-valued columns
    cls = np.repeat(np.arange(1, n_classes + 1), n_per_class).astype(np.int32)
    return raw, cls


def _run(seed, n_per_class, n_classes, n_features, n_test):
    import torch
    from wndcharm import autograd
    lib = _lib()
    raw, cls = _space(seed, n_per_class, n_classes, n_features)
    model = autograd.WNDModel(raw, cls, n_classes, used_fraction=0.3)
    rng = np.random.default_rng(seed + 1)
    x = torch.from_numpy(raw[rng.integers(0, len(raw), n_test)] + rng.normal(0, 0.5, (n_test, n_features))).cuda()
    want = model._forward(x)
    got = torch.full_like(want, float('nan'))
    rc = lib.wndgpu_ad_wnd_forward_noalloc(wndcharm._context(), *model._args(), autograd._vp(x), n_test,
                                           autograd._vp(got), autograd._stream(x))
    assert rc == 0, lib.wndgpu_last_error(wndcharm._context())
    torch.cuda.synchronize()
    return got.cpu().numpy(), want.cpu().numpy(), model, x


def test_noalloc_matches_work_buffer_forward_twice():
    print('no-alloc WND5 forward: call 1 (240 samples x 700 features, 64 test vectors)', flush=True)
    got1, want1, _, _ = _run(11, 60, 4, 700, 64)
    assert np.array_equal(got1, want1)
    assert np.allclose(got1.sum(1), 1.0)
    print('no-alloc WND5 forward: call 2 (90 samples x 300 features, 5 test vectors)', flush=True)
    got2, want2, _, _ = _run(12, 30, 3, 300, 5)
    assert np.array_equal(got2, want2)
    assert np.allclose(got2.sum(1), 1.0)


def test_noalloc_refuses_sizes_above_the_array():
    from wndcharm import autograd
    lib = _lib()
    lib.wndgpu_last_error.restype = ctypes.c_char_p
    lib.wndgpu_last_error.argtypes = [_P]
    _, _, model, x = _run(13, 10, 2, 50, 3)
    per_test = 8 * (model.n_features + model.n_samples + model.n_classes + 2)
    too_many = (32 << 20) // per_test + 1
    # refused on the size check, before any pointer is used
    rc = lib.wndgpu_ad_wnd_forward_noalloc(wndcharm._context(), *model._args(), autograd._vp(x), too_many,
                                           autograd._vp(x), autograd._stream(x))
    assert rc == -1
    assert b'above the' in lib.wndgpu_last_error(wndcharm._context())
    # the largest accepted size still runs
    import torch
    n = (32 << 20) // per_test
    xs = x[:1].repeat(n, 1).contiguous()
    prob = torch.empty((n, model.n_classes), dtype=torch.float64, device=x.device)
    rc = lib.wndgpu_ad_wnd_forward_noalloc(wndcharm._context(), *model._args(), autograd._vp(xs), n,
                                           autograd._vp(prob), autograd._stream(xs))
    assert rc == 0
    want = model._forward(xs[:2].contiguous())
    torch.cuda.synchronize()
    assert torch.equal(prob[:2], want) and torch.equal(prob[-1], want[0])
