"""The allocation-free adjoint entry point (mmsb_stgd_grad_f32_noalloc, scratch in a fixed device
array of the library) against the workspace entry point it shares its kernels with
(mmsb_stgd_grad_f32): equal loss and gradient, bit for bit, on two calls in a row -- the second
smaller, over the scratch the first left behind -- and sizes beyond the scratch rejected."""
import ctypes

import numpy as np
import pytest

from oracle import core

pytestmark = pytest.mark.gpu


def _inputs(dev, k, nb, seed):
    import torch
    rng = np.random.default_rng(seed)
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)
    ys = [t(rng.standard_normal((nb, k)), torch.float32) for _ in range(3)]
    mus = [t(0.3 * rng.standard_normal(k), torch.float64) for _ in range(3)]
    phi = t(rng.standard_normal((k, k)), torch.float64)
    return ys, mus, phi


def _noalloc(mmsb, ys, mus, phi, coeffs, theta):
    import torch
    nb, k = ys[0].shape
    grad = torch.full((k, k), np.nan, dtype=torch.float64, device=ys[0].device)
    loss = torch.full((1,), np.nan, dtype=torch.float64, device=ys[0].device)
    co = np.ascontiguousarray(coeffs, np.float64)
    mmsb.check(mmsb.lib.mmsb_stgd_grad_f32_noalloc(None, k, nb, *[mmsb.ptr(y) for y in ys], ys[0].stride(0),
                                                   *[mmsb.ptr(m) for m in mus], mmsb.ptr(phi), co.ctypes.data,
                                                   float(theta), mmsb.ptr(grad), mmsb.ptr(loss)))
    torch.cuda.synchronize()
    return loss, grad


def test_noalloc_gradient_matches_workspace_entry_point(mmsb, dev):
    co = core.gpu_coeffs(1.3)
    for k, nb, seed in ((100, 1024, 1), (24, 300, 2)):      # the second call smaller than the first
        print(f"case k={k} nb={nb}", flush=True)
        ys, mus, phi = _inputs(dev, k, nb, seed)
        l0, g0 = mmsb.stgd_grad(*ys, *mus, phi, co, 2.0)
        l1, g1 = _noalloc(mmsb, ys, mus, phi, co, 2.0)
        assert torch_equal(l0, l1) and torch_equal(g0, g1), (k, nb)
        # and the FP64 oracle's adjoint on the same inputs
        y64 = [y.double().cpu().numpy() for y in ys]
        m64 = [m.cpu().numpy() for m in mus]
        want = core.tensor_grad(phi.cpu().numpy(), *y64, *m64, co, 2.0)
        np.testing.assert_allclose(g1.cpu().numpy(), want, rtol=1e-9, atol=1e-9 * np.abs(want).max())


def torch_equal(a, b):
    import torch
    return bool(torch.equal(a, b))


def test_noalloc_rejects_sizes_beyond_the_scratch(mmsb, dev):
    ys, mus, phi = _inputs(dev, 256, 2217, 3)     # needs more than the 32 MiB scratch
    with pytest.raises(mmsb.MMSBError) as e:
        _noalloc(mmsb, ys, mus, phi, core.gpu_coeffs(1.0), 1.0)
    assert e.value.code == mmsb.MMSB_ERR_ARG and "scratch" in str(e.value)
    ys, mus, phi = _inputs(dev, 256, 2216, 3)     # the largest minibatch that fits at k = 256
    l0, g0 = mmsb.stgd_grad(*ys, *mus, phi, core.gpu_coeffs(1.0), 1.0)
    l1, g1 = _noalloc(mmsb, ys, mus, phi, core.gpu_coeffs(1.0), 1.0)
    assert torch_equal(g0, g1) and torch_equal(l0, l1)
