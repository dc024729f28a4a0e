"""The pipeline at the sizes of the benchmark configurations, through the C-ABI:

* 100k nodes per partition, k = 100 (alpha0 = 1, P = 0.01 / 2e-4, minibatch 1024): whitening, neighbour
  vectors, STGD steps and Pi against the FP64 oracle at full size, and the adjoint at k = 100 against
  central finite differences, the oracle adjoint and the reference's analytic gradient.
* 1M nodes per partition, k = 500, degree-corrected (the STGD throughput workload): properties that do
  not need an n^2 oracle -- W^T M2 W = I through the factored M2, STGD steps against the oracle on the
  first minibatches, Phi orthonormal after an epoch, Pi-hat rows on the simplex, and Pi and the
  memberships against the oracle on a 10k-node sample.
* 64 uneven instances advanced by one batch, one member against the oracle.
* 4M nodes per partition, k = 1000, checkpointed STGD (minibatch rows rebuilt from CSR): steps against
  the oracle on the first minibatches and Pi on a 10k-node sample.

Tolerances: FP32 storage of tall matrices with FP64 factorisations (1e-4 of the column scale, looser
where the spectrum is nearly degenerate, as stated at each check); FP32 STGD 1e-4 of Phi; BF16 STGD
the update (Phi - Phi0) to 2e-2 - 1e-1, as stated at each check.
"""
import math
import time

import numpy as np
import pytest

from oracle import core

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def subspace_sin(A, B):
    """Largest principal-angle sine between the column spans of A and B."""
    qa, _ = np.linalg.qr(A)
    qb, _ = np.linalg.qr(B)
    s = np.linalg.svd(qa.T @ qb, compute_uv=False)
    return float(np.sqrt(max(0.0, 1.0 - s.min() ** 2)))


def _blocks(g):
    return {nm: core.Csr(g[nm][0], g[nm][1], g[nm][3], g[nm][4]) for nm in ("XA", "XB", "XC", "BA", "CA")}


def test_c2_full_size_matches_oracle(mmsb):
    n, k, b = 100000, 100, 1024
    t0 = time.time()
    g = mmsb.generate(n, n, n, n, k, 1.0, 0.01, 2e-4, seed=0, threads=16)
    print(f"case c2: generated {g['n_edges']} edges in {time.time() - t0:.1f}s", flush=True)
    cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, alpha_0=1.0, learn_rate=1e-3, minibatch=b, oversample=10,
                      power_iters=1, seed=1)
    p = mmsb.Pipeline(cfg)
    p.set_graph(g)
    p.whiten()
    bl = _blocks(g)
    del g
    ref = core.whiten(bl["XA"], bl["XB"], bl["XC"], core.WhitenConfig(k=k, alpha0=1.0, oversample=10, power_iters=1,
                                                                        seed=1))
    ys = core.yvectors(bl["XA"], bl["XB"], bl["XC"], ref)
    print(f"oracle whitening done at {time.time() - t0:.1f}s", flush=True)
    for nm in ("mu_a", "mu_b", "mu_c"):
        np.testing.assert_allclose(p.get(nm), ref["mu_" + nm[-1].upper()], rtol=1e-15)
    np.testing.assert_allclose(p.get("sigma"), ref["sigma"], rtol=1e-5)
    np.testing.assert_allclose(p.get("lam_m2"), ref["lam_M2"], rtol=1e-5)
    # At this size the 99 eigenvalues of M2 below the first lie within 13% of each other (gaps ~1e-3
    # relative), so single eigenvectors are ill-determined while the k-dimensional spans are not:
    # spans of W, U, V and Y to 1e-4, the vectors themselves column by column to a looser bound.
    W = p.get("W")
    for nm, key in (("W", "W"), ("U", "U"), ("V", "V")):
        got = p.get(nm)
        sin = subspace_sin(got, ref[key])
        print(f"{nm}: subspace sine {sin:.2e}, rel err {rel_err(got, ref[key]):.2e}", flush=True)
        assert sin < 1e-4, nm
        assert rel_err(got, ref[key]) < 2e-2, nm
    for nm, key in (("YA", "Y_A"), ("YB", "Y_B"), ("YC", "Y_C")):
        got = p.get(nm)
        assert subspace_sin(got, ys[key]) < 1e-4, nm
        assert rel_err(got, ys[key]) < 2e-2, nm
    # W^T M2 W = I with the oracle's M2 operator (exact products on the oracle's factors)
    WtMW = W.T @ ref["M2"](W)
    assert np.abs(WtMW - np.eye(k)).max() < 1e-3
    # STGD: 20 steps from the seeded Phi0 on the library's own vectors (the oracle steps the same inputs)
    ya, yb, yc = (p.get(nm) for nm in ("YA", "YB", "YC"))
    mus = [p.get(nm) for nm in ("muw_a", "muw_b", "muw_c")]
    phi0 = core.init_phi(1, k)
    p.set_phi(phi0)
    p.stgd(20, "fp32")
    phi_ref, log_ref = core.stgd(ya, yb, yc, *mus, phi0, alpha0=1.0, eta=1e-3, batch=b, epochs=1, steps=20)
    phi = p.get("phi")
    assert rel_err(phi, phi_ref) < 1e-4
    np.testing.assert_allclose(p.get("frob"), log_ref, rtol=1e-3)
    p.eigen()
    p.post()
    _, lam_p, _ = core.eigenvalues(phi)
    np.testing.assert_allclose(p.get("lambda_p"), lam_p, rtol=1e-6)
    pis = core.post_proc(ref, phi, lam_p, bl["XA"], bl["BA"], bl["CA"])
    for nm, P in zip(("pi_x", "pi_a", "pi_b", "pi_c"), pis):
        assert rel_err(p.get(nm), P) < 2e-2, nm             # through W, U: the spans above
    np.testing.assert_allclose(p.get("pi_x"), ya @ (phi * p.get("lambda_p")), rtol=1e-4, atol=1e-4 * np.abs(
        p.get("pi_x")).max())                                # Pi_X = Y_A Phi Lambda' on the library's own Y
    # the adjoint at k = 100 on the first minibatch
    import torch
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    sl = slice(0, b)
    co = core.gpu_coeffs(1.0)
    y64 = [t(y[sl]) for y in (ya, yb, yc)]
    m64 = [t(mu) for mu in mus]
    phi_t = t(phi)
    loss, grad = mmsb.stgd_grad(*y64, *m64, phi_t, co, 1.0)
    want = core.tensor_grad(phi, ya[sl], yb[sl], yc[sl], *mus, co, 1.0)
    np.testing.assert_allclose(grad.cpu().numpy(), want, rtol=1e-9, atol=1e-10 * np.abs(want).max())
    rng = np.random.default_rng(0)
    h = 1e-6
    for _ in range(12):
        i, j = rng.integers(0, k, 2)
        e = np.zeros((k, k))
        e[i, j] = h
        lp, _ = mmsb.stgd_grad(*y64, *m64, t(phi + e), co, 1.0)
        lm, _ = mmsb.stgd_grad(*y64, *m64, t(phi - e), co, 1.0)
        fd = (float(lp[0]) - float(lm[0])) / (2 * h)
        assert abs(fd - float(grad[i, j])) < 1e-5 * max(1.0, abs(float(grad[i, j]))), (i, j, fd, float(grad[i, j]))
    # symmetric single-node inputs: the reference's analytic gradient (Util.cpp:257-304) with theta = 1e4
    y = ya[5]
    mu = mus[0]
    y1 = t(np.stack([y]))                                  # one row, a fresh C-ordered array
    _, g1 = mmsb.stgd_grad(y1, y1, y1, t(mu), t(mu), t(mu), phi_t, core.paper_coeffs(1.0), 1e4)
    ref_g = core.ref_diff_loss_vectors(phi, y, y, y, mu, mu, mu, 1.0, theta=1e4)
    np.testing.assert_allclose(g1.cpu().numpy(), ref_g, rtol=1e-9, atol=1e-9 * np.abs(ref_g).max())
    p.close()


def _c3_graph(mmsb, n=1000000, k=500):
    N = 4 * n
    p_out = 80.0 / (2.0 * N)
    return mmsb.generate(n, n, n, n, k, 2.0, (k + 1) * p_out, p_out, seed=0, degree_corrected=True, threads=16)


def test_c3_full_size_properties(mmsb):
    n, k, b = 1000000, 500, 1024
    t0 = time.time()
    g = _c3_graph(mmsb, n, k)
    print(f"case c3: generated {g['n_edges']} edges in {time.time() - t0:.1f}s", flush=True)
    cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, alpha_0=2.0, learn_rate=1e-6, minibatch=b, oversample=10,
                      power_iters=1, seed=1, epochs=1, polish_epochs=0)
    p = mmsb.Pipeline(cfg)
    p.set_graph(g)
    del g
    p.whiten()
    print(f"whitened at {time.time() - t0:.1f}s", flush=True)
    # W^T M2 W = I with M2 applied through the library's own factors (paper scaling, alpha0 = 2):
    # M2 = 3/2 (L_C S L_B^T + L_B S L_C^T) - 2 mu_A mu_A^T
    W = p.get_f32("W")
    sig = p.get("sigma")
    mu = p.get("mu_a").astype(np.float32)
    A = W.T @ p.get_f32("LC")
    B = W.T @ p.get_f32("LB")
    wm = W.T @ mu
    del W
    WtMW = 1.5 * ((A * sig) @ B.T + (B * sig) @ A.T) - 2.0 * np.outer(wm, wm)
    dev_i = float(np.abs(WtMW - np.eye(k)).max())
    print(f"|W^T M2 W - I|_max = {dev_i:.2e}", flush=True)
    assert dev_i < 2e-3
    # 32 FP32 steps from the seeded Phi0 against the oracle on the same rows
    p.stgd(0, "fp32")
    phi0 = p.get("phi")
    S = 32
    ya, yb, yc = (p.get_rows(nm, 0, S * b) for nm in ("YA", "YB", "YC"))
    mus = [p.get(nm) for nm in ("muw_a", "muw_b", "muw_c")]
    p.stgd(S, "fp32")
    phi_ref, _ = core.stgd(ya, yb, yc, *mus, phi0, alpha0=2.0, eta=1e-6, batch=b, epochs=1, steps=S)
    del ya, yb, yc
    phi = p.get("phi")
    print(f"STGD {S} steps: rel err {rel_err(phi, phi_ref):.2e}, moved {rel_err(phi_ref, phi0):.2e}", flush=True)
    assert rel_err(phi - phi0, phi_ref - phi0) < 1e-2      # the update itself, FP32 against FP64
    assert rel_err(phi, phi_ref) < 1e-4
    # a whole BF16 epoch (the rest of the first, from the step the FP32 run stopped at)
    p.stgd(math.ceil(n / b) - S, "bf16")
    phi = p.get("phi")
    # the renormalisation's k^3 products run as split BF16 (about 16 significant bits) in BF16 mode
    assert np.abs(phi.T @ phi - np.eye(k)).max() < 1e-4
    p.eigen()
    p.post()
    lam_p = p.get("lambda_p")
    # Pi-hat rows: on the simplex (or all zero), non-negative; memberships = the positive entries
    s0, ns = n // 2, 10000
    pin = p.get_rows("pin_x", s0, ns)
    sums = pin.sum(1)
    assert (pin >= 0).all()
    assert np.all((np.abs(sums - 1.0) < 1e-5) | (sums == 0.0))
    np.testing.assert_array_equal(p.get_rows("count_x", s0, ns), (pin > 0).sum(1))
    # Pi and Pi-hat on a 10k sample against the oracle's Y_A Phi Lambda' (library rows and Phi)
    ys = p.get_rows("YA", s0, ns)
    _, lam_p64, _ = core.eigenvalues(phi)
    np.testing.assert_allclose(lam_p, lam_p64, rtol=1e-6)
    pi64 = ys @ (phi * lam_p64)
    pi = p.get_rows("pi_x", s0, ns)
    assert rel_err(pi, pi64) < 1e-4
    pin64 = core.norm_prob_rows(pi64)
    # away from ties (|pos| and |neg| within 1e-4 of each other, entries within 1e-6 of zero) the
    # memberships are the oracle's exactly
    pos = np.where(pi64 > 0, pi64, 0).sum(1)
    neg = -np.where(pi64 > 0, 0, pi64).sum(1)
    clear = np.abs(pos - neg) > 1e-4 * (pos + neg)
    sure = clear[:, None] & (np.abs(pi64) > 1e-6 * np.abs(pi64).max(1, keepdims=True))
    assert np.array_equal((pin > 0)[sure], (pin64 > 0)[sure])
    assert rel_err(pin[clear], pin64[clear]) < 1e-3
    print(f"c3 sample: {clear.mean():.4f} of the nodes away from a sign tie, total {time.time() - t0:.1f}s", flush=True)
    p.close()


def test_c4_batch_of_64_uneven_instances(mmsb):
    """The batched workload: 64 instances, n log-uniform in [5k, 200k], k in {10, 20, 50, 100, 200},
    alpha0 in [0.5, 3], P = 0.02 / 5e-4 scaled by min(1, 5k / n) (at the stated P the mean degree grows
    to ~2000 at 200k nodes, ~8e8 edges for one instance; scaled, every instance has the degree of a
    5k-node one), instance i seeded by i; one BF16 epoch of every instance by one batch; the smallest
    and the largest member against the oracle's epoch on the same whitened vectors."""
    rng = np.random.default_rng(2024)
    count, b = 64, 512
    specs = []
    for i in range(count):
        n = int(round(math.exp(rng.uniform(math.log(5000), math.log(200000)))))
        specs.append((n, int(rng.choice([10, 20, 50, 100, 200])), float(rng.uniform(0.5, 3.0))))
    t0 = time.time()
    pipes = []
    for i, (n, k, a0) in enumerate(specs):
        sc = min(1.0, 5000.0 / n)
        g = mmsb.generate(n, n, n, n, k, a0, 0.02 * sc, 5e-4 * sc, seed=i, threads=16)
        cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, alpha_0=a0, learn_rate=1e-4, minibatch=b, oversample=10,
                          power_iters=1, seed=i + 1, epochs=1, polish_epochs=0)
        p = mmsb.Pipeline(cfg)
        p.set_graph(g)
        del g
        p.whiten()
        p.stgd(0, "fp32")
        pipes.append(p)
        if i % 8 == 7:
            print(f"c4: {i + 1} instances whitened at {time.time() - t0:.1f}s", flush=True)
    print(f"case c4: {count} instances whitened in {time.time() - t0:.1f}s", flush=True)
    sizes = [n for n, _, _ in specs]
    pick = [int(np.argmin(sizes)), int(np.argmax(sizes))]
    before = {i: (pipes[i].get("phi"), [pipes[i].get(nm) for nm in ("YA", "YB", "YC")],
                  [pipes[i].get(nm) for nm in ("muw_a", "muw_b", "muw_c")]) for i in pick}
    batch = mmsb.Batch(pipes, "bf16")
    batch.run(1)
    batch.close()
    for i in range(count):
        assert pipes[i].get("frob").size == -(-specs[i][0] // b)
    for i in pick:
        n, k, a0 = specs[i]
        phi0, ys, mus = before[i]
        phi_ref, _ = core.stgd(*ys, *mus, phi0, alpha0=a0, eta=1e-4, batch=b, epochs=1)
        err = rel_err(pipes[i].get("phi") - phi0, phi_ref - phi0)
        print(f"member {i} (n={n}, k={k}): update rel err {err:.2e}", flush=True)
        assert err < 2e-2
    for p in pipes:
        p.close()


def test_c5_checkpointed_full_size(mmsb):
    """4M nodes per partition, k = 1000, 40 nonzeros per row in each X block (480M over the three):
    the resident neighbour vectors would not fit, so the minibatch rows are rebuilt from CSR each
    step. 32 FP32 and BF16 steps against the oracle on the rows the library serves, and Pi on a
    10k-node sample."""
    n, k, b = 4000000, 1000, 1024
    N = 4 * n
    p_out = 160.0 / (2.0 * N)
    t0 = time.time()
    print("case c5: generating", flush=True)
    g = mmsb.generate(n, n, n, n, k, 1.0, (k + 1) * p_out, p_out, seed=0, threads=16, verbose=True)
    print(f"case c5: generated {g['n_edges']} edges in {time.time() - t0:.1f}s", flush=True)
    cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, alpha_0=1.0, learn_rate=1e-5, minibatch=b, oversample=10,
                      power_iters=1, seed=1, epochs=1, polish_epochs=0)
    p = mmsb.Pipeline(cfg)
    p.set_graph(g)
    del g
    p.whiten()
    assert mmsb.lib.mmsb_get_size(p.ctx, mmsb.Q["V"]) == 0     # recompute form chosen (V released)
    print(f"whitened at {time.time() - t0:.1f}s", flush=True)
    p.stgd(0, "fp32")
    phi0 = p.get("phi")
    S = 32
    ya, yb, yc = (p.get_rows(nm, 0, S * b) for nm in ("YA", "YB", "YC"))
    mus = [p.get(nm) for nm in ("muw_a", "muw_b", "muw_c")]
    phi_ref, _ = core.stgd(ya, yb, yc, *mus, phi0, alpha0=1.0, eta=1e-5, batch=b, epochs=1, steps=S)
    del ya, yb, yc
    # Phi moves ~1e-3 of its size over these steps, so the error of the update itself -- FP32 or BF16
    # contractions against FP64 -- is checked alongside that of Phi
    for prec, tol_upd, tol_phi in (("fp32", 1e-2, 1e-5), ("bf16", 1e-1, 1e-4)):
        p.set_phi(phi0)
        p.stgd(S, prec)
        phi = p.get("phi")
        err = rel_err(phi - phi0, phi_ref - phi0)
        print(f"c5 {prec}: update rel err {err:.2e}, phi rel err {rel_err(phi, phi_ref):.2e}, moved "
              f"{rel_err(phi_ref, phi0):.2e}", flush=True)
        assert err < tol_upd and rel_err(phi, phi_ref) < tol_phi
        p.stgd(math.ceil(n / b) - S, prec)                   # back to an epoch boundary
    p.set_phi(phi_ref)
    p.eigen()
    p.post()
    s0, ns = n // 2, 10000
    ys = p.get_rows("YA", s0, ns)
    _, lam_p64, _ = core.eigenvalues(phi_ref)
    pi64 = ys @ (phi_ref * lam_p64)
    assert rel_err(p.get_rows("pi_x", s0, ns), pi64) < 1e-4
    print(f"c5 done in {time.time() - t0:.1f}s", flush=True)
    p.close()


def test_recovers_planted_communities_at_scale(mmsb):
    """End to end at 250k nodes per partition (1M nodes, 157M edges), k = 50, alpha0 = 1, 60 expected
    neighbours per partition inside a node's community and 20 outside (P = 0.012 / 8e-5): BF16 STGD
    with minibatch 1024 at eta = 3e-2 for 8 epochs, FP32 polish epoch last; the normalised memberships
    of X, matched one-to-one to the generator's Dirichlet memberships, correlate at >= 0.9 on average,
    and the BF16 epochs with the FP32 polish epoch end on the Phi that nine FP32 epochs reach
    (the FP64 oracle reaches 0.976 on a 20k-node graph of the same degrees; the power method, the
    alternative solver, is held to the same bound). At the benchmark sizes the community eigenvalues of
    the pair moments lie at their noise level (about 40 expected neighbours inside the community against
    40-80 outside), where neither the oracle nor the device recovers the communities."""
    from scipy.optimize import linear_sum_assignment
    import importlib
    ev = importlib.import_module("mmsb_amd.evaluate")
    n, k = 250000, 50
    t0 = time.time()
    g = mmsb.generate(n, n, n, n, k, 1.0, 60.0 / (n / k), 20.0 / n, seed=0, threads=16, keep_truth=True)
    print(f"case recovery: {g['n_edges']} edges in {time.time() - t0:.1f}s", flush=True)
    x = np.nonzero(g["part"] == 0)[0]
    truth = np.zeros((n, k))
    truth[g["local"][x]] = g["pi"][x]
    for key in ("pi", "theta", "part", "local"):
        del g[key]
    cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, alpha_0=1.0, learn_rate=3e-2, minibatch=1024, oversample=10,
                      power_iters=1, seed=1, epochs=8, polish_epochs=1, precision="bf16")
    p = mmsb.Pipeline(cfg)
    p.set_graph(g)
    del g
    p.whiten()
    p.stgd(0, "fp32")
    phi0 = p.get("phi")

    def matched():
        p.eigen()
        p.post()
        r = np.nan_to_num(ev.correlations(p.get("pin_x"), truth))
        ri, ci = linear_sum_assignment(-r)
        return float(r[ri, ci].mean())
    start = matched()
    p.stgd(-1, "bf16")
    got = matched()
    phi_mixed = p.get("phi")
    # the same nine epochs in FP32 from the same Phi0: with a constant step the iteration settles on a
    # fixed point, and the FP32 polishing epoch carries the BF16 trajectory onto FP32's (columns matched
    # one to one, with signs: which column lands on which eigenvector depends on the trajectory)
    p.set_phi(phi0)
    p.stgd(9 * -(-n // 1024), "fp32")
    phi32 = p.get("phi")
    c = phi_mixed.T @ phi32
    ri, ci = linear_sum_assignment(-np.abs(c))
    perm = np.empty(k, int)
    perm[ci] = ri
    aligned = phi_mixed[:, perm] * np.sign(c[perm, np.arange(k)])
    gap = float(np.abs(aligned - phi32).max() / np.abs(phi32).max())
    p.set_phi(phi0)
    p.power_method(30)
    pm = matched()
    print(f"matched correlation: Phi0 {start:.3f}, STGD {got:.3f}, power method {pm:.3f}; BF16 + polish vs FP32 "
          f"{gap:.2e} ({time.time() - t0:.1f}s)", flush=True)
    assert start < 0.5 and got > 0.9 and pm > 0.9
    assert gap < 1e-4
    p.close()
