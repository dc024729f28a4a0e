"""GPU parity of the pipeline stages (through the C-ABI) against the FP64 oracle.

Tolerances: tall whitening outputs are FP32 in HBM (FP64 Grams and factorisations), so they agree
with the FP64 oracle to ~1e-4 relative to the column scale; STGD runs FP32 MFMA contractions and is
compared step-for-step from identical inputs at 1e-4; integer outputs (means x n_X, membership
counts away from the threshold) are exact.
"""
import numpy as np
import pytest

from oracle import core, gen

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def make_case(mmsb, n=300, k=4, alpha0=1.0, p_in=0.5, p_out=0.02, seed=0, power_iters=2, oversample=10,
              minibatch=100, **cfg_extra):
    g = mmsb.generate(n, n, n, n, k, alpha0, p_in, p_out, seed=seed)
    cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, alpha_0=alpha0, oversample=oversample,
                      power_iters=power_iters, seed=seed + 1, minibatch=minibatch, **cfg_extra)
    bl = {nm: core.Csr(g[nm][0], g[nm][1], g[nm][3], g[nm][4]) for nm in ("XA", "XB", "XC", "BA", "CA")}
    wcfg = core.WhitenConfig(k=k, alpha0=alpha0, oversample=oversample, power_iters=power_iters, seed=seed + 1,
                             m2_scaling="paper" if cfg.m2_scaling == 0 else "gpu")
    return g, cfg, bl, wcfg


# (60, 6000): partitions of >= 4096 rows with r, k >= 32 take the FP64-MFMA tall products, and r >= 64
# takes the block Jacobi kernel
@pytest.mark.parametrize("k,n,scaling", [(4, 300, "paper"), (10, 1000, "paper"), (6, 400, "gpu"), (60, 6000, "gpu")])
def test_whitening_matches_oracle(mmsb, k, n, scaling):
    g, cfg, bl, wcfg = make_case(mmsb, n=n, k=k, m2_scaling=scaling, p_in=0.6 if k > 4 else 0.5,
                                 p_out=0.02 if n < 4096 else 0.004)
    print(f"case k={k} n={n} scaling={scaling}", flush=True)
    p = mmsb.Pipeline(cfg)
    p.set_graph(g)
    p.whiten()
    ref = core.whiten(bl["XA"], bl["XB"], bl["XC"], wcfg)
    ys = core.yvectors(bl["XA"], bl["XB"], bl["XC"], ref)
    # column means: exact ratios of integers
    for nm in ("mu_a", "mu_b", "mu_c"):
        np.testing.assert_allclose(p.get(nm), ref["mu_" + nm[-1].upper()], rtol=1e-15)
    np.testing.assert_allclose(p.get("sigma"), ref["sigma"], rtol=1e-5)
    assert rel_err(p.get("U"), ref["U"]) < 1e-4
    assert rel_err(p.get("V"), ref["V"]) < 1e-4
    np.testing.assert_allclose(p.get("lam_m2"), ref["lam_M2"], rtol=1e-4)
    assert rel_err(p.get("W"), ref["W"]) < 1e-4
    for nm, key in (("YA", "Y_A"), ("YB", "Y_B"), ("YC", "Y_C")):
        assert rel_err(p.get(nm), ys[key]) < 1e-4, nm
    for nm, key in (("muw_a", "mu_A"), ("muw_b", "mu_B"), ("muw_c", "mu_C")):
        assert rel_err(p.get(nm), ys[key]) < 1e-4, nm
    # the defining invariant on the oracle's exact dense M2: W^T M2 W = I
    we = core.whiten_exact(bl["XA"], bl["XB"], bl["XC"], wcfg)
    W = p.get("W")
    assert np.abs(W.T @ we["M2"] @ W - np.eye(k)).max() < 1e-3
    sweeps = p.get("sweeps")              # both r x r Jacobi factorisations converged before the cap
    assert sweeps.shape == (2,) and (sweeps >= 1).all() and (sweeps < 40).all(), sweeps
    p.close()


# b = 1 is the reference's one-node step (stpm.c:497-790); b = 1500 > n_X makes every step the whole
# partition (one step per epoch, the rows beyond n_X zero)
@pytest.mark.parametrize("k,b,renorm", [(10, 100, 1), (10, 64, 1), (20, 256, 1), (12, 128, 0), (70, 200, 1),
                                        (10, 1, 1), (12, 1500, 1)])
def test_stgd_steps_match_oracle(mmsb, k, b, renorm):
    n = 1000
    rng = np.random.default_rng(k + b)
    ya, yb, yc = (rng.standard_normal((n, k)).astype(np.float32).astype(np.float64) for _ in range(3))
    mua, mub, muc = (0.3 * rng.standard_normal(k) for _ in range(3))
    phi0 = core.init_phi(7, k)
    cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, alpha_0=1.3, learn_rate=2e-3, minibatch=b,
                      renorm_every=renorm, theta=1.0, oversample=0)
    print(f"case k={k} b={b} renorm={renorm}", flush=True)
    p = mmsb.Pipeline(cfg)
    p.set_yvectors(ya, yb, yc, mua, mub, muc)
    p.set_phi(phi0)
    steps = 25
    p.stgd(steps, "fp32")
    phi_ref, log_ref = core.stgd(ya, yb, yc, mua, mub, muc, phi0, alpha0=1.3, eta=2e-3, batch=b, epochs=1,
                                 renorm_every=renorm, steps=steps)
    assert rel_err(p.get("phi"), phi_ref) < 1e-4
    np.testing.assert_allclose(p.get("frob"), log_ref, rtol=2e-3, atol=1e-12)
    p.close()


@pytest.mark.parametrize("k,b,iters", [(10, 100, 1), (70, 200, 1), (20, 256, 2)])
def test_stgd_quintic_matches_oracle(mmsb, k, b, iters):
    """Degree-5 Newton-Schulz renormalisation (Gram, polynomial and update kernels) step for step
    against the oracle's degree-5 iteration."""
    n = 1000
    rng = np.random.default_rng(k * b + iters)
    ya, yb, yc = (rng.standard_normal((n, k)).astype(np.float32).astype(np.float64) for _ in range(3))
    mua, mub, muc = (0.3 * rng.standard_normal(k) for _ in range(3))
    phi0 = core.init_phi(9, k)
    cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, alpha_0=1.1, learn_rate=2e-3, minibatch=b, ns_iters=iters,
                      ns_degree=5, oversample=0)
    print(f"case k={k} b={b} iters={iters}", flush=True)
    p = mmsb.Pipeline(cfg)
    p.set_yvectors(ya, yb, yc, mua, mub, muc)
    p.set_phi(phi0)
    steps = 25
    p.stgd(steps, "fp32")
    phi_ref, log_ref = core.stgd(ya, yb, yc, mua, mub, muc, phi0, alpha0=1.1, eta=2e-3, batch=b, epochs=1,
                                 steps=steps, ns_iters=iters, ns_degree=5)
    assert rel_err(p.get("phi"), phi_ref) < 1e-4
    np.testing.assert_allclose(p.get("frob"), log_ref, rtol=2e-3, atol=1e-12)
    p.close()


@pytest.mark.parametrize("degree,iters", [(3, 2), (5, 1)])
def test_stgd_guarded_renorm_matches_oracle(mmsb, degree, iters):
    """Three rows of 12x norm (hub nodes) push ||Phi~^T Phi~||_inf to 3-6, so the scale guard of the
    Newton-Schulz iteration fires in most of the steps; the kernels follow the oracle's guarded
    iteration (the unguarded one diverges from such starts)."""
    n, k, b = 600, 16, 100
    rng = np.random.default_rng(42)
    ys = [rng.standard_normal((n, k)).astype(np.float32).astype(np.float64) for _ in range(3)]
    for y in ys:
        y[[50, 170, 333]] *= 12.0
    mus = [0.2 * rng.standard_normal(k) for _ in range(3)]
    phi0 = core.init_phi(5, k)
    cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, alpha_0=1.0, learn_rate=5e-3, minibatch=b, ns_iters=iters,
                      ns_degree=degree, oversample=0)
    print(f"case degree={degree} iters={iters}", flush=True)
    p = mmsb.Pipeline(cfg)
    p.set_yvectors(*ys, *mus)
    p.set_phi(phi0)
    p.stgd(12, "fp32")
    phi_ref, log_ref = core.stgd(*ys, *mus, phi0, alpha0=1.0, eta=5e-3, batch=b, epochs=1, steps=12, ns_iters=iters,
                                 ns_degree=degree)
    assert rel_err(p.get("phi"), phi_ref) < 1e-4
    np.testing.assert_allclose(p.get("frob"), log_ref, rtol=2e-3, atol=1e-12)
    # the guard's row sums are order-independent (fixed-point atomics): repeated runs give Phi bit for
    # bit, in both precisions, though the tiles' atomics land in a different order each time; the log
    # entries (sums of per-tile partials in a fixed order) repeat bit for bit as well
    for prec in ("fp32", "bf16"):
        outs, logs = [], []
        for _ in range(3):
            q = mmsb.Pipeline(cfg)
            q.set_yvectors(*ys, *mus)
            q.set_phi(phi0)
            q.stgd(12, prec)
            outs.append(q.get("phi"))
            logs.append(q.get("frob"))
            q.close()
        assert all(np.array_equal(outs[0], o) for o in outs[1:]), prec
        assert all(np.array_equal(logs[0], f) for f in logs[1:]), prec
    p.close()


def test_stgd_full_epochs_and_wraparound(mmsb):
    """Several epochs over a batch size that does not divide n_X (last batch partial)."""
    n, k, b = 1037, 8, 100
    rng = np.random.default_rng(3)
    ya, yb, yc = (rng.standard_normal((n, k)).astype(np.float32).astype(np.float64) for _ in range(3))
    mus = [0.2 * rng.standard_normal(k) for _ in range(3)]
    phi0 = core.init_phi(1, k)
    cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, alpha_0=0.7, learn_rate=1e-3, minibatch=b, epochs=3,
                      oversample=0)
    p = mmsb.Pipeline(cfg)
    p.set_yvectors(ya, yb, yc, *mus)
    p.set_phi(phi0)
    p.stgd(-1, "fp32")
    phi_ref, log_ref = core.stgd(ya, yb, yc, *mus, phi0, alpha0=0.7, eta=1e-3, batch=b, epochs=3)
    assert len(p.get("frob")) == len(log_ref) == 3 * 11
    assert rel_err(p.get("phi"), phi_ref) < 1e-4
    p.close()


def test_seeded_phi0_matches_oracle(mmsb):
    n, k = 256, 9
    rng = np.random.default_rng(0)
    ys = [rng.standard_normal((n, k)) for _ in range(3)]
    cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, seed=11, oversample=0, minibatch=64)
    p = mmsb.Pipeline(cfg)
    p.set_yvectors(*ys, np.zeros(k), np.zeros(k), np.zeros(k))
    p.stgd(0, "fp32")
    np.testing.assert_allclose(p.get("phi"), core.init_phi(11, k), atol=1e-6)
    p.close()


def test_eigen_post_match_oracle(mmsb):
    g, cfg, bl, wcfg = make_case(mmsb, n=400, k=5, p_in=0.6, p_out=0.02, seed=2)
    p = mmsb.Pipeline(cfg)
    p.set_graph(g)
    p.whiten()
    ref = core.whiten(bl["XA"], bl["XB"], bl["XC"], wcfg)
    # a non-orthonormal Phi exercises lambda = ||phi||^3 and Lambda' = ||lambda|| / lambda
    rng = np.random.default_rng(9)
    phi = np.linalg.qr(rng.standard_normal((5, 5)))[0] * np.array([1.0, 1.1, 0.9, 1.3, 0.8])
    p.set_phi(phi)
    p.eigen()
    lam, lam_p, ah = core.eigenvalues(phi)
    np.testing.assert_allclose(p.get("lambda"), lam, rtol=1e-6)
    np.testing.assert_allclose(p.get("lambda_p"), lam_p, rtol=1e-6)
    np.testing.assert_allclose(p.get("alpha_hat"), ah, rtol=1e-6)
    p.post()
    pis = core.post_proc(ref, phi, lam_p, bl["XA"], bl["BA"], bl["CA"])
    for nm, P in zip(("pi_x", "pi_a", "pi_b", "pi_c"), pis):
        assert rel_err(p.get(nm), P) < 2e-4, nm
        Pn = core.norm_prob_rows(P)
        got = p.get("pin_" + nm[-1])
        assert np.abs(got - Pn).max() < 2e-4, nm
        # membership counts agree wherever no entry sits within 1e-3 of the threshold (tau = 0)
        cnt = p.get("count_" + nm[-1])
        want = (Pn > 0).sum(1)
        clear = ~((Pn > 0) & (Pn < 1e-3)).any(1)
        assert np.array_equal(cnt[clear], want[clear]), nm
    p.close()


def test_read_inputs_and_write_outputs(mmsb, tmp_path):
    """The drop-in file path: FOLDER_NAME/DBLP_Gx_a.txt ... in, Pi_* / evecs / evals / frob / sim_info out."""
    from oracle import io as oio
    n, k = 200, 3
    g = mmsb.generate(n, n, n, n, k, 1.0, 0.5, 0.03, seed=4)
    for nm, stem in (("XA", "Gx_a"), ("XB", "Gx_b"), ("XC", "Gx_c"), ("BA", "Gb_a"), ("CA", "Gc_a")):
        oio.write_coo(str(tmp_path / f"DBLP_{stem}.txt"), g[nm][0], g[nm][1])
    cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, folder_name=str(tmp_path) + "/", minibatch=64, epochs=2,
                      alpha_0_str="1.0", learn_rate_str="1e-3")
    p = mmsb.Pipeline(cfg)
    p.read_inputs()
    p.whiten()
    p.stgd(-1)
    p.eigen()
    p.post()
    p.write_outputs()
    for stem in ("Pi_X", "Pi_A", "Pi_B", "Pi_C", "evecs", "evals", "frob_norm", "sim_info"):
        assert (tmp_path / f"{stem}_{k}_1.0_1e-3.txt").exists(), stem
    vals = np.loadtxt(tmp_path / f"Pi_X_{k}_1.0_1e-3.txt")
    np.testing.assert_allclose(vals.reshape(n, k), p.get("pi_x"), rtol=1e-6, atol=1e-9)
    ev = np.loadtxt(tmp_path / f"evecs_{k}_1.0_1e-3.txt")
    np.testing.assert_allclose(ev.reshape(k, k).T, p.get("phi"), rtol=1e-6)   # column-major
    assert np.isclose(np.loadtxt(tmp_path / f"evals_{k}_1.0_1e-3.txt").sum(), 1.0)
    p.close()


@pytest.mark.parametrize("k,b,degree", [(10, 128, 3), (70, 200, 3), (24, 64, 3), (70, 200, 5), (10, 1, 3)])
def test_stgd_bf16_tracks_oracle(mmsb, k, b, degree):
    """BF16 contractions (FP32 accumulation, split-BF16 k^3 products): Phi within 5e-3 of the FP64
    oracle after 25 steps, and the FP32 polish epoch brings it back to the FP32 trajectory's level."""
    n = 1000
    rng = np.random.default_rng(100 + k)
    ya, yb, yc = (rng.standard_normal((n, k)).astype(np.float32).astype(np.float64) for _ in range(3))
    mus = [0.3 * rng.standard_normal(k) for _ in range(3)]
    phi0 = core.init_phi(3, k)
    iters = 2 if degree == 3 else 1
    cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, alpha_0=1.3, learn_rate=2e-3, minibatch=b, oversample=0,
                      ns_iters=iters, ns_degree=degree)
    print(f"case k={k} b={b} degree={degree}", flush=True)
    p = mmsb.Pipeline(cfg)
    p.set_yvectors(ya, yb, yc, *mus)
    p.set_phi(phi0)
    p.stgd(25, "bf16")
    phi_ref, log_ref = core.stgd(ya, yb, yc, *mus, phi0, alpha0=1.3, eta=2e-3, batch=b, epochs=1, steps=25,
                                 ns_iters=iters, ns_degree=degree)
    got = p.get("phi")
    assert rel_err(got, phi_ref) < 5e-3
    assert np.abs(got.T @ got - np.eye(k)).max() < 1e-4      # split-BF16 renormalisation keeps Phi orthonormal
    assert rel_err(p.get("frob"), log_ref) < 0.05
    p.close()


def test_get_rows_matches_get(mmsb):
    n, k = 300, 5
    rng = np.random.default_rng(1)
    ys = [rng.standard_normal((n, k)) for _ in range(3)]
    cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, oversample=0, minibatch=64)
    p = mmsb.Pipeline(cfg)
    p.set_yvectors(*ys, np.zeros(k), np.zeros(k), np.zeros(k))
    for nm, y in zip(("YA", "YB", "YC"), ys):
        np.testing.assert_allclose(p.get_rows(nm, 37, 100), y[37:137], rtol=1e-6)
    p.close()


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_epoch_graph_replays_match_eager(mmsb, prec):
    """Captured epoch graphs, replayed several times with eager work and host syncs in between,
    give bit-identical Phi to direct launches after every epoch (odd and even steps per epoch)."""
    for n in (1024, 960):                         # 16 and 15 steps per epoch at b = 64
        k, b = 16, 64
        rng = np.random.default_rng(n)
        ys = [rng.standard_normal((n, k)) for _ in range(3)]
        mus = [0.2 * rng.standard_normal(k) for _ in range(3)]
        phi0 = core.init_phi(5, k)
        ps = []
        for use in (1, 0):
            cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, oversample=0, minibatch=b, learn_rate=3e-3,
                              use_graphs=use)
            p = mmsb.Pipeline(cfg)
            p.set_yvectors(*ys, *mus)
            p.set_phi(phi0)
            ps.append(p)
        spe = n // b
        for epoch in range(4):
            for p in ps:
                if epoch == 2:
                    p.set_learn_rate(1e-3)      # drops the captured graphs; later epochs recapture
                p.stgd(spe, prec)
                p.eigen()                       # eager work between replays
            a, e = ps[0].get("phi"), ps[1].get("phi")
            assert np.array_equal(a, e), f"n={n} epoch {epoch}"
        np.testing.assert_allclose(ps[0].get("frob"), ps[1].get("frob"), rtol=1e-12)
        for p in ps:
            p.close()


@pytest.mark.parametrize("orth", [True, False])
def test_power_method_matches_oracle(mmsb, orth):
    """Batch implicit power method (CPU tensorDecom_alpha0 with the orthogonalisation added)."""
    n, k = 700, 6
    rng = np.random.default_rng(21)
    ys = [rng.standard_normal((n, k)).astype(np.float32).astype(np.float64) for _ in range(3)]
    mus = [0.3 * rng.standard_normal(k) for _ in range(3)]
    phi0 = core.init_phi(2, k)
    cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, alpha_0=0.8, oversample=0, minibatch=64)
    p = mmsb.Pipeline(cfg)
    p.set_yvectors(*ys, *mus)
    p.set_phi(phi0)
    lam = p.power_method(5, orthogonalize=orth)
    phi_ref, lam_ref = core.power_method(*ys, *mus, phi0, alpha0=0.8, iters=5, orthogonalize=orth)
    np.testing.assert_allclose(lam, lam_ref, rtol=1e-5)
    assert rel_err(p.get("phi"), phi_ref) < 1e-5
    p.close()


def test_power_method_recovers_odeco(mmsb):
    from tests.test_oracle import odeco_data
    V, Y = odeco_data(k=4, n=2000, seed=3)
    z = np.zeros(4)
    rng = np.random.default_rng(4)
    phi0 = np.linalg.qr(V + 0.2 * rng.standard_normal((4, 4)))[0]
    cfg = mmsb.config(nx=2000, na=2000, nb=2000, nc=2000, ncom=4, alpha_0=1e-9, oversample=0, minibatch=64)
    p = mmsb.Pipeline(cfg)
    p.set_yvectors(Y, Y, Y, z, z, z)
    p.set_phi(phi0)
    p.power_method(30)
    assert np.allclose(np.sort(np.abs(V.T @ p.get("phi")).max(0)), 1.0, atol=1e-5)
    p.close()


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_timed_steps_match_untimed(mmsb, prec):
    """mmsb_stgd_timed runs the same steps as mmsb_stgd (bit-identical Phi) and reports positive
    per-phase device times."""
    n, k, b = 900, 24, 64
    rng = np.random.default_rng(8)
    ys = [rng.standard_normal((n, k)) for _ in range(3)]
    mus = [0.2 * rng.standard_normal(k) for _ in range(3)]
    phi0 = core.init_phi(4, k)
    ps = []
    for _ in range(2):
        cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, oversample=0, minibatch=b, learn_rate=2e-3, use_graphs=0)
        p = mmsb.Pipeline(cfg)
        p.set_yvectors(*ys, *mus)
        p.set_phi(phi0)
        ps.append(p)
    ps[0].stgd(20, prec)
    ms = ps[1].stgd_timed(20, prec)
    assert all(t > 0 for t in ms), ms
    assert np.array_equal(ps[0].get("phi"), ps[1].get("phi"))
    for p in ps:
        p.close()


@pytest.mark.parametrize("prec,degree", [("fp32", 3), ("bf16", 3), ("bf16", 5)])
def test_batch_matches_single_instances(mmsb, prec, degree):
    """Uneven instances (n, k, minibatch, alpha0, eta all differ; odd and even steps per epoch)
    advanced together by the persistent batch launch give the same Phi, bit for bit, as each
    instance run alone, and the same ||dPhi||^2 log; also after autotuning the grid."""
    specs = [(700, 8, 64, 0.7, 2e-3), (1300, 24, 128, 1.5, 1e-3), (520, 70, 64, 2.5, 3e-3), (2000, 10, 200, 1.0, 2e-3)]
    singles, members = [], []
    for idx, (n, k, b, a0, eta) in enumerate(specs):
        rng = np.random.default_rng(idx)
        ys = [rng.standard_normal((n, k)) for _ in range(3)]
        mus = [0.2 * rng.standard_normal(k) for _ in range(3)]
        phi0 = core.init_phi(idx + 1, k)
        for lst in (singles, members):
            cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, alpha_0=a0, oversample=0, minibatch=b, learn_rate=eta,
                              ns_iters=2 if degree == 3 else 1, ns_degree=degree)
            p = mmsb.Pipeline(cfg)
            p.set_yvectors(*ys, *mus)
            p.set_phi(phi0)
            lst.append(p)
    print(f"case prec={prec} degree={degree} instances={len(specs)}", flush=True)
    batch = mmsb.Batch(members, prec)
    batch.run(2)
    grid, ms = batch.autotune()
    assert grid >= 0 and ms > 0          # 0: grouped per-slot launches won
    batch.run(1)
    for idx, ((n, k, b, a0, eta), p_single, p_batch) in enumerate(zip(specs, singles, members)):
        spe = -(-n // b)
        p_single.stgd(3 * spe, prec)
        assert np.array_equal(p_single.get("phi"), p_batch.get("phi")), (n, k, b)
        np.testing.assert_allclose(p_batch.get("frob"), p_single.get("frob"), rtol=1e-9)
        # and every member against the FP64 oracle's three epochs on the same inputs
        rng = np.random.default_rng(idx)
        ys = [rng.standard_normal((n, k)) for _ in range(3)]
        mus = [0.2 * rng.standard_normal(k) for _ in range(3)]
        phi_ref, log_ref = core.stgd(*ys, *mus, core.init_phi(idx + 1, k), alpha0=a0, eta=eta, batch=b, epochs=3,
                                     ns_iters=2 if degree == 3 else 1, ns_degree=degree)
        assert rel_err(p_batch.get("phi"), phi_ref) < (1e-4 if prec == "fp32" else 2e-2), (n, k, b)
    batch.close()
    for p in singles + members:
        p.close()


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_c1_recovers_planted_communities(mmsb, prec):
    """The whole pipeline on the 1000-node-per-partition, k = 10 workload (alpha0 = 1, P = 0.6 / 0.02,
    minibatch 100, eta = 1e-3, 200 epochs): the normalised memberships of the X partition are scored
    against the generator's Dirichlet memberships with the p-value test of CPU/Pvalue.cpp."""
    import importlib
    ev = importlib.import_module("mmsb_amd.evaluate")
    n, k = 1000, 10
    g = mmsb.generate(n, n, n, n, k, 1.0, 0.6, 0.02, seed=0, keep_truth=True)
    cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, alpha_0=1.0, learn_rate=1e-3, minibatch=100, epochs=200,
                      polish_epochs=1 if prec == "bf16" else 0, oversample=10, power_iters=2, seed=1, precision=prec)
    print(f"case prec={prec}", flush=True)
    p = mmsb.Pipeline(cfg)
    p.set_graph(g)
    p.whiten()
    p.stgd(-1, prec)
    p.eigen()
    p.post()
    x_nodes = np.nonzero(g["part"] == 0)[0]
    truth = np.zeros((n, k))
    truth[g["local"][x_nodes]] = g["pi"][x_nodes]
    res = ev.evaluate(p.get("pin_x"), truth)
    best = res["correlations"].max(0)                 # strongest positive match of each true community
    print(f"recovery {res['recovery_ratio']:.2f} error {res['error']:.4f} matches {res['matches']} "
          f"best correlations min {best.min():.3f} median {np.median(best):.3f}", flush=True)
    assert res["recovery_ratio"] == 1.0 and res["error"] < 0.2
    p.close()


def test_stop_request_ends_stgd_at_an_epoch_boundary(mmsb):
    """forceExit_flag (stpm.c:61,754,1198): a stop request ends the STGD epochs at an epoch boundary
    and returns normally -- requested before the call (one epoch runs), and from another thread while
    the epochs run. Phi is that of the same number of uninterrupted steps."""
    import threading
    n, k, b = 4096, 32, 256
    spe = n // b
    rng = np.random.default_rng(2)
    ys = [rng.standard_normal((n, k)) for _ in range(3)]
    mus = [0.1 * rng.standard_normal(k) for _ in range(3)]
    phi0 = core.init_phi(3, k)
    cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, oversample=0, minibatch=b, epochs=5000, learn_rate=1e-4,
                      sgd_iter_min=0, sgd_iter_max=0)
    p = mmsb.Pipeline(cfg)
    p.set_yvectors(*ys, *mus)
    p.set_phi(phi0)
    p.request_stop()
    p.stgd(-1, "fp32")
    assert p.get("frob").size == spe
    timer = threading.Timer(0.05, p.request_stop)
    timer.start()
    p.stgd(-1, "fp32")
    timer.join()
    done = p.get("frob").size
    print(f"stopped after {done // spe} epochs", flush=True)
    assert done % spe == 0 and 2 * spe <= done < 5001 * spe
    q = mmsb.Pipeline(cfg)
    q.set_yvectors(*ys, *mus)
    q.set_phi(phi0)
    q.stgd(done, "fp32")
    assert np.array_equal(p.get("phi"), q.get("phi"))
    p.close()
    q.close()


def test_cli_matches_the_api(mmsb, tmp_path):
    """./algo GPU_A (the reference's main, stpm.c:1068-1218: pre_proc, GPU_Task on a pthread,
    post_proc) on partition files writes the same Pi / evecs / evals files, byte for byte, as the
    same configuration through the context API (frob_norm to its FP64 atomic summation order)."""
    import os
    import subprocess
    from oracle import io as oio
    n, k = 200, 3
    g = mmsb.generate(n, n, n, n, k, 1.0, 0.5, 0.03, seed=4)
    api, cli = tmp_path / "api", tmp_path / "cli"
    for d in (api, cli):
        d.mkdir()
        for nm, stem in (("XA", "Gx_a"), ("XB", "Gx_b"), ("XC", "Gx_c"), ("BA", "Gb_a"), ("CA", "Gc_a")):
            oio.write_coo(str(d / f"DBLP_{stem}.txt"), g[nm][0], g[nm][1])
    cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, folder_name=str(api) + "/", minibatch=64, epochs=2,
                      alpha_0_str="1.0", learn_rate_str="1e-3")
    p = mmsb.Pipeline(cfg)
    p.read_inputs()
    p.whiten()
    p.stgd(-1)
    p.eigen()
    p.post()
    p.write_outputs()
    p.close()
    exe = os.path.join(os.path.dirname(mmsb.__file__), "algo")
    r = subprocess.run([exe, "GPU_A", f"NX={n}", f"NA={n}", f"NB={n}", f"NC={n}", f"NCOM={k}", f"FOLDER_NAME={cli}/",
                        "MINIBATCH=64", "EPOCHS=2", "ALPHA_0=1.0", "LEARN_RATE=1e-3"],
                       capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stderr
    assert "Task completed." in r.stdout
    for stem in ("Pi_X", "Pi_A", "Pi_B", "Pi_C", "evecs", "evals"):
        name = f"{stem}_{k}_1.0_1e-3.txt"
        assert (cli / name).read_bytes() == (api / name).read_bytes(), stem
    name = f"frob_norm_{k}_1.0_1e-3.txt"
    np.testing.assert_allclose(np.loadtxt(cli / name), np.loadtxt(api / name), rtol=1e-12)
    assert (cli / f"sim_info_{k}_1.0_1e-3.txt").exists()


_CHILD_BATCH = r'''
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import __graft_entry__ as ge
from oracle import core
m = ge.load_package()
specs = [(700, 8, 64, 0.7, 2e-3), (1300, 24, 128, 1.5, 1e-3)]
singles, members = [], []
for idx, (n, k, b, a0, eta) in enumerate(specs):
    rng = np.random.default_rng(idx)
    ys = [rng.standard_normal((n, k)) for _ in range(3)]
    mus = [0.2 * rng.standard_normal(k) for _ in range(3)]
    for lst in (singles, members):
        p = m.Pipeline(m.config(nx=n, na=n, nb=n, nc=n, ncom=k, alpha_0=a0, oversample=0, minibatch=b, learn_rate=eta))
        p.set_yvectors(*ys, *mus)
        p.set_phi(core.init_phi(idx + 1, k))
        lst.append(p)
batch = m.Batch(members, "bf16")
batch.run(1)
grid, ms = batch.autotune()
batch.run(1)
for (n, k, b, a0, eta), ps, pb in zip(specs, singles, members):
    ps.stgd(2 * -(-n // b), "bf16")
    assert np.array_equal(ps.get("phi"), pb.get("phi")), n
print("child batch ok, autotuned grid", grid, flush=True)
'''


def test_batch_in_a_second_process(mmsb, tmp_path):
    """A batch in a second process on the same GPU, after this process has run a persistent
    (cooperative) batch launch: the child's cooperative launch may be refused (only one process can
    hold the global-wave-sync resource); its epochs then run as grouped launches, with the same Phi
    as single instances."""
    import os
    import subprocess
    import sys
    n, k = 640, 8
    rng = np.random.default_rng(5)
    p = mmsb.Pipeline(mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, oversample=0, minibatch=64))
    p.set_yvectors(*[rng.standard_normal((n, k)) for _ in range(3)], *[np.zeros(k)] * 3)
    batch = mmsb.Batch([p], "fp32")
    batch.run(1)                                   # persistent launch in this process
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    script = tmp_path / "child_batch.py"
    script.write_text(_CHILD_BATCH)
    r = subprocess.run([sys.executable, str(script), root], capture_output=True, text=True, timeout=110, cwd=root)
    print(r.stdout.strip(), flush=True)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "child batch ok" in r.stdout
    batch.close()
    p.close()


@pytest.mark.parametrize("mu_scale", [0.005, 0.05])
def test_factorisations_at_large_k(mmsb, mu_scale):
    """k = 1100: the seeded Phi0 (CholeskyQR of a k x k Gaussian: the triangular inverse above 1024
    columns) and the power method's orthonormalisation (the one-sided Jacobi SVD of Phi itself, with
    rotations in 4-column blocks, since 8-column blocks of A and V exceed the LDS at this size),
    against the oracle. With means ~ 0.05 the shared mean terms make the power step's columns nearly
    collinear (cond(Phi) ~ 1e7): the polar factor from Phi's own SVD stays accurate, where one formed
    from Phi^T Phi (cond ~ 1e14) would not."""
    n, k = 1536, 1100
    rng = np.random.default_rng(7)
    ys = [(rng.standard_normal((n, k)) / np.sqrt(k)).astype(np.float32).astype(np.float64) for _ in range(3)]
    mus = [mu_scale * rng.standard_normal(k) for _ in range(3)]
    cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, alpha_0=0.8, oversample=0, minibatch=256, seed=13)
    p = mmsb.Pipeline(cfg)
    p.set_yvectors(*ys, *mus)
    p.stgd(0, "fp32")                                   # the seeded Phi0
    phi0 = core.init_phi(13, k)
    assert rel_err(p.get("phi"), phi0) < 1e-5
    p.set_phi(phi0)
    lam = p.power_method(1, orthogonalize=True)
    phi_ref, lam_ref = core.power_method(*ys, *mus, phi0, alpha0=0.8, iters=1, orthogonalize=True)
    np.testing.assert_allclose(lam, lam_ref, rtol=1e-5)
    got = p.get("phi")
    assert rel_err(got, phi_ref) < 1e-5
    assert np.abs(got.T @ got - np.eye(k)).max() < 1e-6
    p.close()


def test_unequal_partitions_match_oracle(mmsb):
    """Partitions of four different sizes (n_X, n_A, n_B, n_C as the NX..NC macros of stpm.h:47-106
    allow): whitening, the neighbour vectors, STGD steps and Pi of every partition against the oracle."""
    nx, na, nb, nc, k = 900, 700, 800, 650, 6
    g = mmsb.generate(nx, na, nb, nc, k, 1.0, 0.6, 0.02, seed=3)
    cfg = mmsb.config(nx=nx, na=na, nb=nb, nc=nc, ncom=k, alpha_0=1.0, oversample=10, power_iters=2, seed=4,
                      minibatch=128, learn_rate=2e-3)
    bl = {nm: core.Csr(g[nm][0], g[nm][1], g[nm][3], g[nm][4]) for nm in ("XA", "XB", "XC", "BA", "CA")}
    assert (bl["XA"].n_rows, bl["XA"].n_cols, bl["BA"].n_rows, bl["CA"].n_rows) == (nx, na, nb, nc)
    wcfg = core.WhitenConfig(k=k, alpha0=1.0, oversample=10, power_iters=2, seed=4)
    p = mmsb.Pipeline(cfg)
    p.set_graph(g)
    p.whiten()
    ref = core.whiten(bl["XA"], bl["XB"], bl["XC"], wcfg)
    ys = core.yvectors(bl["XA"], bl["XB"], bl["XC"], ref)
    for nm, n in (("mu_a", na), ("mu_b", nb), ("mu_c", nc)):
        got = p.get(nm)
        assert got.shape == (n,), (nm, got.shape)
        np.testing.assert_allclose(got, ref["mu_" + nm[-1].upper()], rtol=1e-15)
    assert rel_err(p.get("W"), ref["W"]) < 1e-4
    for nm, key in (("YA", "Y_A"), ("YB", "Y_B"), ("YC", "Y_C")):
        got = p.get(nm)
        assert got.shape == ys[key].shape == (nx, k), (nm, got.shape)
        assert rel_err(got, ys[key]) < 1e-4, nm
    # STGD from the seeded Phi0 on the device's own vectors (the oracle steps the same inputs)
    ya, yb, yc = (p.get(nm).astype(np.float64) for nm in ("YA", "YB", "YC"))
    mus = [p.get(nm) for nm in ("muw_a", "muw_b", "muw_c")]
    phi0 = core.init_phi(4, k)
    p.set_phi(phi0)
    p.stgd(10, "fp32")
    phi_ref, _ = core.stgd(ya, yb, yc, *mus, phi0, alpha0=1.0, eta=2e-3, batch=128, epochs=1, steps=10)
    phi = p.get("phi")
    assert rel_err(phi, phi_ref) < 1e-4
    p.eigen()
    p.post()
    pis = core.post_proc(ref, phi, p.get("lambda_p"), bl["XA"], bl["BA"], bl["CA"])
    for nm, P, n in zip(("pi_x", "pi_a", "pi_b", "pi_c"), pis, (nx, na, nb, nc)):
        got = p.get(nm)
        assert got.shape == P.shape and P.shape[0] == n, (nm, got.shape, P.shape)
        assert rel_err(got, P) < 2e-4, nm
    p.close()
