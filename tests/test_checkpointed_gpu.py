"""Checkpointed STGD (y_mode "recompute"): each step rebuilds its minibatch's whitened rows from CSR
(y = G_X* rows times W, Z_B^T W, Z_C^T W) instead of reading resident Y_A, Y_B, Y_C^T
(pre_proc.c:339-345 forms them all up front). The rows are summed in the order of the whole-matrix
SpMM and rounded to BF16 the same way, so the two forms must agree bit for bit: W, every Phi, Pi and
the adjoint (the ||dPhi||^2 log to its FP64 summation order). The resident form itself is checked against the FP64 oracle in
test_pipeline_gpu.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def build_pair(mmsb, n, k, b, epochs=1, polish=0, seed=0, p_in=0.5, p_out=0.02):
    g = mmsb.generate(n, n, n, n, k, 1.0, p_in, p_out, seed=seed)
    out = []
    for mode in ("stored", "recompute"):
        cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, alpha_0=1.0, oversample=10, power_iters=1, seed=seed + 1,
                          minibatch=b, learn_rate=2e-3, epochs=epochs, polish_epochs=polish, y_mode=mode)
        p = mmsb.Pipeline(cfg)
        p.set_graph(g)
        p.whiten()
        out.append(p)
    return out


# (k, n, b): r/4 chunks per row below 64 lanes and above (k = 130: two float4 chunks on some lanes),
# a minibatch that does not divide n_X (partial last step) and one that does
@pytest.mark.parametrize("k,n,b", [(12, 1000, 96), (70, 1536, 128), (130, 2200, 256)])
@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_recomputed_minibatches_match_resident(mmsb, k, n, b, prec):
    print(f"case k={k} n={n} b={b} prec={prec}", flush=True)
    res, win = build_pair(mmsb, n, k, b, epochs=2, polish=1 if prec == "bf16" else 0, p_in=0.6)
    assert np.array_equal(res.get("W"), win.get("W"))
    for nm in ("muw_a", "muw_b", "muw_c"):
        assert np.array_equal(res.get(nm), win.get(nm)), nm
    # rows served on demand in the recompute form equal the resident rows
    for nm in ("YA", "YB", "YC"):
        assert np.array_equal(res.get_rows(nm, 37, 200), win.get_rows(nm, 37, 200)), nm
    assert np.array_equal(res.get("YC"), win.get("YC"))
    # two full epochs through captured graphs (BF16 then the FP32 polishing epoch for bf16), then a
    # partial epoch of direct launches
    for p in (res, win):
        p.stgd(-1, prec)
    assert np.array_equal(res.get("phi"), win.get("phi"))
    steps = max(1, (n // b) // 2)
    for p in (res, win):
        p.stgd(steps, prec)
    assert np.array_equal(res.get("phi"), win.get("phi"))
    # ||dPhi||^2 is accumulated by FP64 atomics across tiles (order not fixed)
    np.testing.assert_allclose(res.get("frob"), win.get("frob"), rtol=1e-12)
    # the adjoint reads resident rows in one form and recomputes them in the other
    phi = res.get("phi")
    la, ga = res.grad(phi, 5, 300)
    lb, gb = win.grad(phi, 5, 300)
    assert la == lb and np.array_equal(ga, gb)
    for p in (res, win):
        p.post()
    for nm in ("pi_x", "pi_a", "pi_b", "pi_c", "pin_x", "count_x"):
        assert np.array_equal(res.get(nm), win.get(nm)), nm
    for p in (res, win):
        p.close()


def test_recompute_mode_state_rules(mmsb):
    """What the recompute form gives back after whitening (the sketches and V), what it refuses
    (batches need resident neighbour vectors), and the power method (CPU tensorDecom_alpha0,
    Util.cpp:131-206) over recomputed row chunks, identical to the resident form."""
    res, win = build_pair(mmsb, 600, 8, 64)
    res.get("V")
    with pytest.raises(mmsb.MMSBError):
        win.get("V")
    with pytest.raises(mmsb.MMSBError):
        mmsb.Batch([win], "fp32")
    lam_r, lam_w = res.power_method(3), win.power_method(3)
    assert np.array_equal(lam_r, lam_w) and np.array_equal(res.get("phi"), win.get("phi"))
    # y_mode "auto" keeps the resident form when it fits: V is still there
    g = mmsb.generate(600, 600, 600, 600, 8, 1.0, 0.5, 0.02, seed=0)
    cfg = mmsb.config(nx=600, na=600, nb=600, nc=600, ncom=8, oversample=10, power_iters=1, seed=1, minibatch=64,
                      y_mode="auto")
    p = mmsb.Pipeline(cfg)
    p.set_graph(g)
    p.whiten()
    assert np.array_equal(p.get("V"), res.get("V"))
    for q in (res, win, p):
        q.close()


def test_weighted_blocks_match_oracle_and_recompute(mmsb):
    """EXPECTED-style weighted blocks (a value per nonzero, pre_proc.c:64-65): whitening and the
    neighbour vectors against the FP64 oracle with the values, and the recompute form's weighted
    gather equal to the resident vectors through BF16 STGD."""
    from oracle import core
    n, k = 800, 6
    g = mmsb.generate(n, n, n, n, k, 1.0, 0.6, 0.02, seed=4)
    rng = np.random.default_rng(4)
    for nm in ("XA", "XB", "XC", "BA", "CA"):
        rp, ci, _, nr, nc = g[nm]
        g[nm] = (rp, ci, rng.uniform(0.5, 2.0, ci.size).astype(np.float32), nr, nc)
    bl = {nm: core.Csr(g[nm][0], g[nm][1], g[nm][3], g[nm][4], g[nm][2].astype(np.float64))
          for nm in ("XA", "XB", "XC")}
    wcfg = core.WhitenConfig(k=k, alpha0=1.0, oversample=10, power_iters=1, seed=5)
    ref = core.whiten(bl["XA"], bl["XB"], bl["XC"], wcfg)
    ys = core.yvectors(bl["XA"], bl["XB"], bl["XC"], ref)
    ps = []
    for mode in ("stored", "recompute"):
        cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, alpha_0=1.0, oversample=10, power_iters=1, seed=5,
                          minibatch=128, binary=0, y_mode=mode)
        p = mmsb.Pipeline(cfg)
        p.set_graph(g)
        p.whiten()
        ps.append(p)
    res, win = ps

    def rel_err(a, b):
        return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    assert rel_err(res.get("W"), ref["W"]) < 1e-4
    for nm, key in (("YA", "Y_A"), ("YB", "Y_B"), ("YC", "Y_C")):
        assert rel_err(res.get(nm), ys[key]) < 1e-4, nm
        assert np.array_equal(res.get(nm), win.get(nm)), nm
    for p in ps:
        p.stgd(-1, "bf16")
    assert np.array_equal(res.get("phi"), win.get("phi"))
    for p in ps:
        p.close()


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_recompute_with_unequal_partitions(mmsb, prec):
    """Partitions of four sizes: the recomputed rows gather from blocks of n_A, n_B and n_C columns
    (G_XA W, G_XB Z_B^T W, G_XC Z_C^T W), so each block's own width must be used; the recompute form
    must still equal the resident one bit for bit."""
    nx, na, nb, nc, k, b = 1100, 700, 900, 650, 20, 128
    g = mmsb.generate(nx, na, nb, nc, k, 1.0, 0.6, 0.02, seed=5)
    ps = []
    for mode in ("stored", "recompute"):
        cfg = mmsb.config(nx=nx, na=na, nb=nb, nc=nc, ncom=k, alpha_0=1.0, oversample=10, power_iters=1, seed=6,
                          minibatch=b, learn_rate=2e-3, epochs=1, y_mode=mode)
        p = mmsb.Pipeline(cfg)
        p.set_graph(g)
        p.whiten()
        ps.append(p)
    res, win = ps
    for nm in ("YA", "YB", "YC"):
        assert np.array_equal(res.get(nm), win.get(nm)), nm
    for p in ps:
        p.stgd(-1, prec)
    assert np.array_equal(res.get("phi"), win.get("phi"))
    for p in ps:
        p.post()
    for nm in ("pi_x", "pi_a", "pi_b", "pi_c"):
        assert np.array_equal(res.get(nm), win.get(nm)), nm
    for p in ps:
        p.close()
