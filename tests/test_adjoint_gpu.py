"""The hand-written adjoint kernel (mmsb_stgd_grad_*) behind a torch.autograd.Function (harness
only): finite differences in FP64 (h = 1e-6), the oracle's exact adjoint, and the reference's
analytic gradient (Diff_Loss / The_second_term, CPU Util.cpp:257-304) on symmetric inputs."""
import numpy as np
import pytest

from oracle import core

pytestmark = pytest.mark.gpu


def make_tensor_loss(mmsb):
    import torch

    class TensorLoss(torch.autograd.Function):
        """L(Phi) of the implicit tensor objective; backward returns the kernel's adjoint."""

        @staticmethod
        def forward(ctx, phi, ya, yb, yc, mua, mub, muc, coeffs, theta):
            loss, grad = mmsb.stgd_grad(ya, yb, yc, mua, mub, muc, phi.contiguous(), coeffs, theta)
            ctx.save_for_backward(grad)
            return loss[0].clone()

        @staticmethod
        def backward(ctx, gout):
            (grad,) = ctx.saved_tensors
            return (gout * grad, None, None, None, None, None, None, None, None)

    return TensorLoss


def inputs(dev, k, nb, seed, symmetric=False, dtype=None):
    import torch
    dtype = dtype or torch.float64
    rng = np.random.default_rng(seed)
    ya = rng.standard_normal((nb, k))
    yb = ya.copy() if symmetric else rng.standard_normal((nb, k))
    yc = ya.copy() if symmetric else rng.standard_normal((nb, k))
    mua = rng.standard_normal(k)
    mub = mua.copy() if symmetric else rng.standard_normal(k)
    muc = mua.copy() if symmetric else rng.standard_normal(k)
    phi = rng.standard_normal((k, k)) * 0.5
    t = lambda a, dt=torch.float64: torch.tensor(a, dtype=dt, device=dev)
    return (dict(ya=ya, yb=yb, yc=yc, mua=mua, mub=mub, muc=muc, phi=phi),
            [t(ya, dtype), t(yb, dtype), t(yc, dtype), t(mua), t(mub), t(muc)], t(phi))


@pytest.mark.parametrize("k,nb,scaling", [(4, 9, "gpu"), (12, 50, "paper"), (33, 130, "gpu")])
def test_adjoint_matches_finite_differences_and_oracle(mmsb, dev, k, nb, scaling):
    import torch
    TensorLoss = make_tensor_loss(mmsb)
    h, dev_in, phi = inputs(dev, k, nb, seed=k)
    co = mmsb.tensor_coeffs(1.3, scaling)
    theta = 0.7
    phi.requires_grad_(True)
    loss = TensorLoss.apply(phi, *dev_in, co, theta)
    loss.backward()
    g = phi.grad.detach().cpu().numpy()
    ref_g = core.tensor_grad(h["phi"], h["ya"], h["yb"], h["yc"], h["mua"], h["mub"], h["muc"], tuple(co), theta)
    ref_l = core.tensor_loss(h["phi"], h["ya"], h["yb"], h["yc"], h["mua"], h["mub"], h["muc"], tuple(co), theta)
    np.testing.assert_allclose(float(loss.detach()), ref_l, rtol=1e-11)
    np.testing.assert_allclose(g, ref_g, rtol=1e-10, atol=1e-10 * np.abs(ref_g).max())
    # central finite differences of the GPU loss itself (FP64, h = 1e-6)
    step = 1e-6
    fd = np.zeros_like(g)
    rng = np.random.default_rng(0)
    idx = [(i, j) for i in range(k) for j in range(k)]
    pick = idx if len(idx) <= 64 else [idx[t] for t in rng.choice(len(idx), 64, replace=False)]
    base = phi.detach().clone()
    for (i, j) in pick:
        lp = float(TensorLoss.apply(base + step * _unit(base, i, j), *dev_in, co, theta))
        lm = float(TensorLoss.apply(base - step * _unit(base, i, j), *dev_in, co, theta))
        fd[i, j] = (lp - lm) / (2 * step)
    sel = tuple(np.array(pick).T)
    scale = np.abs(g).max()
    assert np.abs(fd[sel] - g[sel]).max() < 1e-6 * max(scale, 1.0)


def _unit(t, i, j):
    import torch
    e = torch.zeros_like(t)
    e[i, j] = 1.0
    return e


def test_adjoint_equals_reference_gradient_symmetric(mmsb, dev):
    """Single node, y_A = y_B = y_C, mu_A = mu_B = mu_C: the kernel equals the reference's
    3 theta sum <phi_i,phi_j>^2 phi_j - 3 T(phi_i, phi_i, I) with the paper coefficients."""
    k = 8
    h, dev_in, phi = inputs(dev, k, 1, seed=3, symmetric=True)
    co = mmsb.tensor_coeffs(2.0, "paper")
    _, grad = mmsb.stgd_grad(*dev_in, phi, co, 1e4)
    ref = core.ref_diff_loss_vectors(h["phi"], h["ya"][0], h["yb"][0], h["yc"][0], h["mua"], h["mub"], h["muc"], 2.0, 1e4)
    np.testing.assert_allclose(grad.cpu().numpy(), ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())


def test_adjoint_fp32_inputs(mmsb, dev):
    import torch
    k, nb = 20, 256
    h, dev_in, phi = inputs(dev, k, nb, seed=5, dtype=torch.float32)
    co = mmsb.tensor_coeffs(1.0, "gpu")
    loss, grad = mmsb.stgd_grad(*dev_in, phi, co, 1.0)
    y32 = [np.asarray(h[n], np.float32).astype(np.float64) for n in ("ya", "yb", "yc")]
    ref = core.tensor_grad(h["phi"], *y32, h["mua"], h["mub"], h["muc"], tuple(co), 1.0)
    np.testing.assert_allclose(grad.cpu().numpy(), ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())


def test_grad_recomputed_from_csr_equals_stored(mmsb):
    """mmsb_grad with the minibatch rebuilt from CSR (checkpointed form) equals the stored-y form."""
    n, k = 500, 6
    g = mmsb.generate(n, n, n, n, k, 1.0, 0.4, 0.02, seed=8)
    cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, oversample=10, power_iters=2, minibatch=64)
    p = mmsb.Pipeline(cfg)
    p.set_graph(g)
    p.whiten()
    phi = core.init_phi(4, k)
    l1, g1 = p.grad(phi, 123, 200, recompute=False)
    l2, g2 = p.grad(phi, 123, 200, recompute=True)
    assert abs(l1 - l2) <= 1e-6 * abs(l1)
    np.testing.assert_allclose(g2, g1, rtol=1e-5, atol=1e-6 * np.abs(g1).max())
    # and both equal the oracle adjoint on the stored (FP32) rows
    ys = [p.get_rows(nm, 123, 200) for nm in ("YA", "YB", "YC")]
    mus = [p.get(nm) for nm in ("muw_a", "muw_b", "muw_c")]
    ref = core.tensor_grad(phi, *ys, *mus, tuple(mmsb.tensor_coeffs(1.0, "gpu")), 1.0)
    np.testing.assert_allclose(g1, ref, rtol=1e-8, atol=1e-8 * np.abs(ref).max())
    p.close()
