"""Pins the FP64 oracle itself (CPU): known-answer vectors, closed forms, brute force and
invariants, each chosen so that a plausible mistake (dropped term, wrong sign or index,
transposed operand) fails."""
import json
import os

import numpy as np
import pytest

from oracle import core, gen, io, philox

GOLD = os.path.join(os.path.dirname(__file__), "golden")
REF_DATA = "/root/reference/CommunityDetectionSoftware_CPU/data"


# ---- RNG and elementary functions ------------------------------------------------------------
def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32_10
    kats = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
            ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
            ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
             (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]
    for c, k, want in kats:
        got = philox.philox4x32(*[np.uint32(x) for x in c], *[np.uint32(x) for x in k])
        assert tuple(int(g) for g in got) == want


def test_det_log_exp_accuracy():
    x = np.exp(np.random.default_rng(0).uniform(-700, 700, 20000))
    x = np.r_[x, 1.0, 0.5, 2.0, np.nextafter(1.0, 2.0), 1e-300, 1e300]
    rel = np.abs(philox.det_log(x) - np.log(x)) / np.maximum(np.abs(np.log(x)), 1e-300)
    assert rel.max() < 1e-15
    assert philox.det_log(np.array([1.0]))[0] == 0.0
    y = np.random.default_rng(1).uniform(-700, 700, 20000)
    assert (np.abs(philox.det_exp(y) - np.exp(y)) / np.exp(y)).max() < 1e-15
    assert philox.det_exp(np.array([0.0]))[0] == 1.0


def test_uniforms_range():
    w = philox.draw(7, 1, np.arange(100000, dtype=np.uint32), 0)
    u, v = philox.u53(w[0], w[1]), philox.u53_open0(w[2], w[3])
    assert u.min() >= 0 and u.max() < 1 and v.min() > 0 and v.max() <= 1
    assert abs(u.mean() - 0.5) < 0.005


# ---- generator ---------------------------------------------------------------------------------
def test_gamma_moments():
    # Gamma(1+a): mean 1+a, variance 1+a
    a = 0.25
    n = 200000
    g = gen.gamma1p(3, np.arange(n, dtype=np.uint32), np.zeros(n, np.uint32), a)
    assert abs(g.mean() - (1 + a)) < 4 * np.sqrt((1 + a) / n)
    assert abs(g.var() - (1 + a)) < 0.03


def test_dirichlet_moments():
    # Dirichlet(alpha0/k): E pi_c = 1/k, E pi_c^2 = (a+1) a / (a0 (a0+1)) with a = a0/k
    p = gen.GenParams(20000, 0, 0, 0, 10, 1.0, 0.1, 0.01, seed=1)
    pi = gen.gen_pi(p)
    assert np.allclose(pi.sum(1), 1.0, atol=1e-12)
    a0, k = 1.0, 10
    a = a0 / k
    assert np.abs(pi.mean(0) - 1 / k).max() < 0.01
    want2 = (a + 1) * a / (a0 * (a0 + 1))
    assert abs((pi ** 2).mean() - want2) < 0.005


def test_pareto_theta_mean_and_tail():
    p = gen.GenParams(50000, 0, 0, 0, 2, 1.0, 0.1, 0.01, seed=2, degree_corrected=True, pareto_shape=2.5)
    th = gen.gen_theta(p)
    assert abs(th.mean() - 1.0) < 1e-12
    # shape-2.5 Pareto normalised by its mean 5/3: P(theta > t) = (t * 5/3)^-2.5
    t = 2.0
    assert abs((th > t).mean() - (t * 5 / 3) ** -2.5) < 0.01


def test_exact_edges_match_probability():
    p = gen.GenParams(150, 150, 150, 150, 4, 1.0, 0.3, 0.02, seed=4)
    g = gen.generate(p)
    th, pi = g.theta, g.pi
    P = (np.outer(th, th)) * (p.p_out + (p.p_in - p.p_out) * pi @ pi.T)
    iu = np.triu_indices(p.n_total, 1)
    mean = P[iu].sum()
    sd = np.sqrt((P[iu] * (1 - P[iu])).sum())
    assert abs(len(g.edges) - mean) < 5 * sd
    # edges are unique, ordered, no self loops
    assert np.all(g.edges[:, 0] < g.edges[:, 1])
    key = g.edges[:, 0] * p.n_total + g.edges[:, 1]
    assert np.all(np.diff(key) > 0)


def test_ball_counts_and_rates():
    p = gen.GenParams(500, 500, 500, 500, 5, 0.5, 0.04, 0.002, seed=5, mode="balls")
    th = gen.gen_theta(p)
    pi = gen.gen_pi(p)
    cum, m_bg, members, cums, m_c = gen.ball_plan(p, th, pi)
    W = cum[-1]
    assert m_bg == int(np.floor(p.p_out * W * W / 2 + 0.5))
    S = (th[:, None] * pi).sum(0)
    assert np.allclose(m_c, np.floor((p.p_in - p.p_out) * S * S / 2 + 0.5))
    e = gen.edges_balls(p, th, pi)
    P = np.outer(th, th) * (p.p_out + (p.p_in - p.p_out) * pi @ pi.T)
    expect = P[np.triu_indices(p.n_total, 1)].sum()
    assert 0.9 * expect < len(e) < 1.02 * expect   # collisions only remove edges


def test_split_is_a_partition_and_blocks_route_edges():
    p = gen.GenParams(30, 20, 25, 15, 3, 1.0, 0.5, 0.1, seed=6)
    g = gen.generate(p)
    for part, size in enumerate([30, 20, 25, 15]):
        loc = np.sort(g.local[g.part == part])
        assert np.array_equal(loc, np.arange(size))
    # brute force: every X-A edge lands in G_XA at (local X, local A)
    rp, ci, nr, nc = g.blocks["XA"]
    D = core.csr_dense(rp, ci, nr, nc)
    want = np.zeros((nr, nc))
    for i, j in g.edges:
        for u, v in ((i, j), (j, i)):
            if g.part[u] == 0 and g.part[v] == 1:
                want[g.local[u], g.local[v]] = 1
    assert np.array_equal(D, want)


# ---- I/O -----------------------------------------------------------------------------------------
def test_reader_excerpt_orientations(tmp_path):
    path = os.path.join(GOLD, "Gx_a_head3000.txt")
    rows = np.loadtxt(path, dtype=np.int64)
    n = int(rows.max())
    rp, ci, v, _, _ = io.read_coo(path, n, n, first_is_row=True)
    assert v is None and ci.size == 3000
    D = core.csr_dense(rp, ci, n, n)
    assert np.all(D[rows[:, 0] - 1, rows[:, 1] - 1] == 1)
    rpT, ciT, _, _, _ = io.read_coo(path, n, n, first_is_row=False)
    assert np.array_equal(core.csr_dense(rpT, ciT, n, n), D.T)


def test_reader_duplicate_policies(tmp_path):
    f = tmp_path / "w.txt"
    f.write_text("1 2 0.5\n3 1 1.5\n1 2 0.25\n")
    rp, ci, v, _, _ = io.read_coo(str(f), 3, 3, weighted=True, duplicates="sum")
    assert np.allclose(core.csr_dense(rp, ci, 3, 3, v)[0, 1], 0.75)
    rp, ci, v, _, _ = io.read_coo(str(f), 3, 3, weighted=True, duplicates="last")
    assert np.allclose(core.csr_dense(rp, ci, 3, 3, v)[0, 1], 0.25)
    # the while(!feof) repeat of the last entry: summed it doubles the last weight (IOfn.cpp:36-51)
    rp, ci, v, _, _ = io.read_coo(str(f), 3, 3, weighted=True, duplicates="sum", feof_repeat=True)
    assert np.allclose(core.csr_dense(rp, ci, 3, 3, v)[0, 1], 1.0)
    g = tmp_path / "b.txt"
    g.write_text("2 3\n1 1\n")
    rp, ci, v, _, _ = io.read_coo(str(g), 3, 3, feof_repeat=True)
    assert ci.size == 2 and v is None   # binary: the repeat is harmless


@pytest.mark.skipif(not os.path.isdir(REF_DATA), reason="reference data not present")
def test_reader_full_data_matches_digests():
    dig = json.load(open(os.path.join(GOLD, "data_digests.json")))
    import hashlib
    for name, d in dig.items():
        rp, ci, _, _, _ = io.read_coo(os.path.join(REF_DATA, name + ".txt"), d["rows"], d["cols"])
        assert ci.size == d["nnz"] == d["lines"]
        assert hashlib.sha256(rp.astype(np.int32).tobytes()).hexdigest() == d["rowptr_sha256"]
    assert dig["Gx_b"]["file_sha256"] == dig["Gx_c"]["file_sha256"]


def test_writer_formats():
    assert io.format_mat([1.0, -0.5]) == "%5.25e\n%5.25e\n" % (1.0, -0.5)
    assert io.format_mat([0.1]).startswith("1.0000000000000000555111512e-01")
    P = np.array([[0.5, 0.0], [0.0, 1.0 / 3.0]])
    assert io.format_pi_triplets(P) == "1\t1\t0.5\n2\t2\t0.333333\n"


# ---- whitening -----------------------------------------------------------------------------------
def small_graph(seed=0, n=300, k=4):
    p = gen.GenParams(n, n, n, n, k, 1.0, 0.5, 0.02, seed=seed)
    g = gen.generate(p)
    bl = {nm: core.Csr(rp, ci, nr, nc) for nm, (rp, ci, nr, nc) in g.blocks.items()}
    return g, bl


def test_whitening_randomized_matches_exact_and_whitens():
    g, bl = small_graph()
    cfg = core.WhitenConfig(k=4, alpha0=1.0, oversample=10, power_iters=6, seed=1)
    wr = core.whiten(bl["XA"], bl["XB"], bl["XC"], cfg)
    we = core.whiten_exact(bl["XA"], bl["XB"], bl["XC"], cfg)
    # W^T M2 W = I on the exact dense M2 (the defining invariant of the whitening matrix)
    WtMW = wr["W"].T @ we["M2"] @ wr["W"]
    assert np.abs(WtMW - np.eye(4)).max() < 1e-9
    # W W^T is the pseudo-inverse of the rank-k truncation of M2: unique, compare directly
    WWe = we["W"] @ we["W"].T
    assert np.abs(wr["W"] @ wr["W"].T - WWe).max() < 1e-8 * np.abs(WWe).max()
    assert np.allclose(wr["lam_M2"], we["lam_M2"], rtol=1e-9)
    # the signs are canonical, so W itself agrees
    assert np.abs(wr["W"] - we["W"]).max() < 1e-8 * np.abs(we["W"]).max()
    # means are column means (pre_proc.c:154)
    assert np.allclose(wr["mu_A"], bl["XA"].dense().mean(0))


def test_randomized_error_decays_with_power_iterations():
    # range-finder error shrinks by ~(sigma_{k+p+1}/sigma_k)^2 per power iteration (Halko et al. 2011)
    g, bl = small_graph()
    we = core.whiten_exact(bl["XA"], bl["XB"], bl["XC"], core.WhitenConfig(k=4, alpha0=1.0))
    errs = []
    for q in (1, 2, 3):
        wr = core.whiten(bl["XA"], bl["XB"], bl["XC"], core.WhitenConfig(k=4, alpha0=1.0, power_iters=q, seed=1))
        errs.append(np.abs(wr["W"].T @ we["M2"] @ wr["W"] - np.eye(4)).max())
    assert errs[1] < 0.2 * errs[0] and errs[2] < 0.2 * errs[1]


def test_zb_solves_pair_equation():
    # Z_B P_BC = P_AC projected on the top-k right singular space of P_BC
    g, bl = small_graph(seed=2)
    cfg = core.WhitenConfig(k=4, alpha0=1.0, power_iters=2, seed=3)
    we = core.whiten_exact(bl["XA"], bl["XB"], bl["XC"], cfg)
    A, B, C = bl["XA"].dense(), bl["XB"].dense(), bl["XC"].dense()
    nx = A.shape[0]
    P_AC, P_BC = A.T @ C / nx, B.T @ C / nx
    V = we["V"]
    assert np.allclose(we["Z_B"] @ P_BC @ V, P_AC @ V, atol=1e-10)


def test_gaussian_matrix_moments():
    z = core.gaussian_matrix(9, 0, 2000, 50, fp32=False)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    z2 = core.gaussian_matrix(9, 0, 2000, 50, fp32=True)
    assert np.abs(z2 - z).max() < 1e-6 * np.abs(z).max()


def test_cholqr2_orthonormal():
    X = np.random.default_rng(0).standard_normal((500, 30)) @ np.diag(np.logspace(0, -5, 30))
    Q, R = core.cholqr2(X)
    assert np.abs(Q.T @ Q - np.eye(30)).max() < 1e-12
    assert np.allclose(Q @ R, X, atol=1e-12)
    assert np.allclose(np.tril(R, -1), 0)


# ---- STGD ----------------------------------------------------------------------------------------
def odeco_data(k=4, n=4000, seed=0, s=2.0):
    """y_A = y_B = y_C = s * v_h(x), h(x) uniform: E[y (x) y (x) y] = s^3/k sum_i v_i^(x)3."""
    rng = np.random.default_rng(seed)
    V, _ = np.linalg.qr(rng.standard_normal((k, k)))
    h = rng.integers(0, k, n)
    Y = s * V[:, h].T
    return V, Y


def test_t_hat_brute_force():
    rng = np.random.default_rng(1)
    k, b = 3, 5
    YA, YB, YC = (rng.standard_normal((b, k)) for _ in range(3))
    mA, mB, mC = (rng.standard_normal(k) for _ in range(3))
    phi = rng.standard_normal((k, k))
    c3, c1, c0 = core.gpu_coeffs(1.3)
    T = np.zeros((k, k, k))
    for x in range(b):
        T += (c3 * np.einsum("i,j,l->ijl", YA[x], YB[x], YC[x])
              + c1 * (np.einsum("i,j,l->ijl", YA[x], YB[x], mC) + np.einsum("i,j,l->ijl", YA[x], mB, YC[x])
                      + np.einsum("i,j,l->ijl", mA, YB[x], YC[x]))) / b
    T += c0 * np.einsum("i,j,l->ijl", mA, mB, mC)
    want = np.stack([np.einsum("ijl,i,j->l", T, phi[:, i], phi[:, i]) for i in range(k)], 1)
    assert np.allclose(core.t_hat_c(phi, YA, YB, YC, mA, mB, mC, (c3, c1, c0)), want)


def test_gpu_scaling_matches_init_buff_matrices():
    # c_g = phi~^T S_g phi~, c_mu = phi~^T S_mu phi~ with the 4x4 matrices of init_buff (stpm.c:1026-1038)
    a0 = 0.7
    Sg = np.zeros(16); Smu = np.zeros(16)
    Sg[3] = Sg[9] = -a0 / (a0 + 2); Sg[11] = 1
    Smu[1] = 2 * a0 * a0 / ((a0 + 1) * (a0 + 2)); Smu[11] = -a0 / (a0 + 2)
    Sg, Smu = Sg.reshape(4, 4, order="F"), Smu.reshape(4, 4, order="F")
    rng = np.random.default_rng(2)
    ma, mb, a, b = rng.standard_normal(4)
    ph = np.array([ma, mb, a, b])       # rows mu_A, mu_B, y_A, y_B of iter_mat (stpm.c:458-462)
    c3, c1, c0 = core.gpu_coeffs(a0)
    assert np.isclose(ph @ Sg @ ph, c3 * a * b + c1 * (a * mb + ma * b))
    assert np.isclose(ph @ Smu @ ph, c0 * ma * mb + c1 * a * b)


def test_stgd_recovers_odeco_components():
    V, Y = odeco_data()
    k = V.shape[0]
    rng = np.random.default_rng(3)
    phi0, _ = np.linalg.qr(V + 0.3 * rng.standard_normal((k, k)))
    z = np.zeros(k)
    phi, log = core.stgd(Y, Y, Y, z, z, z, phi0, alpha0=1e-12, eta=0.02, batch=200, epochs=40, ns_iters=3)
    # each column aligned with some +-v_i
    C = np.abs(V.T @ phi)
    assert np.allclose(np.sort(C.max(0)), 1.0, atol=1e-6)
    assert np.abs(phi.T @ phi - np.eye(k)).max() < 1e-8
    assert log[-1] < 1e-12


def test_adjoint_finite_differences():
    rng = np.random.default_rng(4)
    k, b = 4, 7
    YA, YB, YC = (rng.standard_normal((b, k)) for _ in range(3))
    mA, mB, mC = (rng.standard_normal(k) for _ in range(3))
    phi = rng.standard_normal((k, k))
    co = core.gpu_coeffs(0.9)
    g = core.tensor_grad(phi, YA, YB, YC, mA, mB, mC, co, 1.7)
    h = 1e-6
    fd = np.zeros_like(phi)
    for i in range(k):
        for j in range(k):
            e = np.zeros_like(phi); e[i, j] = h
            fd[i, j] = (core.tensor_loss(phi + e, YA, YB, YC, mA, mB, mC, co, 1.7)
                        - core.tensor_loss(phi - e, YA, YB, YC, mA, mB, mC, co, 1.7)) / (2 * h)
    assert np.allclose(g, fd, rtol=1e-6, atol=1e-6)


def test_adjoint_equals_reference_gradient_on_symmetric_inputs():
    rng = np.random.default_rng(5)
    k = 5
    y = rng.standard_normal(k)
    mu = rng.standard_normal(k)
    phi = rng.standard_normal((k, k))
    a0 = 1.4
    ref = core.ref_diff_loss_vectors(phi, y, y, y, mu, mu, mu, a0, theta=1e4)
    ours = core.tensor_grad(phi, y[None], y[None], y[None], mu, mu, mu, core.paper_coeffs(a0), 1e4)
    assert np.allclose(ours, ref, rtol=1e-10, atol=1e-8)
    # asymmetric inputs: they differ (the reference keeps only the C-mode term three times)
    y2 = rng.standard_normal(k)
    ref2 = core.ref_diff_loss_vectors(phi, y, y2, y, mu, mu, mu, a0, theta=1e4)
    ours2 = core.tensor_grad(phi, y[None], y2[None], y[None], mu, mu, mu, core.paper_coeffs(a0), 1e4)
    assert not np.allclose(ours2, ref2)


def test_eigenvalues_and_invlamphi_fixture():
    # with per-step orthonormalisation lambda = 1, Lambda' = sqrt(k) I, alpha_hat = 1/k (stpm.c:799-870)
    k = 10
    Q, _ = np.linalg.qr(np.random.default_rng(6).standard_normal((k, k)))
    lam, lam_p, ah = core.eigenvalues(Q)
    assert np.allclose(lam, 1) and np.allclose(lam_p, np.sqrt(k)) and np.allclose(ah, 1 / k)
    # data/INVLAMPHI_WRITE.txt holds Lambda'^... Phi^T of a k=10 run: M M^T / 10 = I
    t = np.loadtxt(os.path.join(GOLD, "INVLAMPHI_WRITE.txt"))
    M = np.zeros((10, 10))
    M[t[:, 0].astype(int) - 1, t[:, 1].astype(int) - 1] = t[:, 2]
    assert np.abs(M @ M.T / 10 - np.eye(10)).max() < 1e-5
    assert np.abs((Q * lam_p).T @ (Q * lam_p) / 10 - np.eye(10)).max() < 1e-12


def test_power_method_odeco():
    V, Y = odeco_data(k=3, n=3000, seed=7)
    rng = np.random.default_rng(8)
    phi0, _ = np.linalg.qr(V + 0.2 * rng.standard_normal((3, 3)))
    z = np.zeros(3)
    phi, lam = core.power_method(Y, Y, Y, z, z, z, phi0, alpha0=1e-12, iters=30)
    assert np.allclose(np.sort(np.abs(V.T @ phi).max(0)), 1.0, atol=1e-8)


# ---- post-processing -------------------------------------------------------------------------------
def test_norm_prob_rows_closed_forms():
    P = np.array([[0.2, 0.6, -0.2, 0.0],      # pos wins: /0.8, clamp
                  [-0.3, -0.5, 0.2, 0.0],     # neg wins: -v/0.8
                  [0.0, 0.0, 0.0, 0.0],       # all zero stays
                  [0.5, -0.5, 0.0, 0.0]])     # tie |pos| == |neg|: divide by pos
    want = np.array([[0.25, 0.75, 0, 0], [0.375, 0.625, 0, 0], [0, 0, 0, 0], [1, 0, 0, 0]])
    assert np.allclose(core.norm_prob_rows(P), want)
    rp, ids = core.memberships(core.norm_prob_rows(P))
    assert rp.tolist() == [0, 2, 4, 4, 5] and ids.tolist() == [0, 1, 0, 1, 0]


def test_polar_renorm_converges_to_polar_factor():
    """Newton-Schulz with the scale guard reaches U V^T of the SVD (the polar factor): from a
    near-orthonormal start (guard inactive, cubic formula unchanged) and from an iterate whose
    largest singular value is far beyond sqrt(3), where the unguarded iteration diverges."""
    rng = np.random.default_rng(5)
    k = 12
    Q, _ = np.linalg.qr(rng.standard_normal((k, k)))
    near = Q + 1e-3 * rng.standard_normal((k, k))
    far = Q @ np.diag(np.geomspace(0.3, 80.0, k)) @ np.linalg.qr(rng.standard_normal((k, k)))[0]
    for X, iters in ((near, 4), (far, 40)):
        U, _, Vt = np.linalg.svd(X)
        got = core.polar_renorm(X, iters)
        assert np.abs(got - U @ Vt).max() < 1e-10
    I = np.eye(k)
    one = near @ (1.5 * I - 0.5 * (near.T @ near))            # plain cubic step: guard inactive
    np.testing.assert_allclose(core.polar_renorm(near, 1), one, rtol=0, atol=1e-15)
    with np.errstate(over="ignore", invalid="ignore"):
        x = far.copy()
        for _ in range(40):
            x = x @ (1.5 * I - 0.5 * x.T @ x)
    assert not np.isfinite(x).all() or np.abs(x).max() > 1e6    # the unguarded iteration blows up


def test_quintic_newton_schulz_maps_singular_values():
    """One degree-5 step acts on the singular values alone: X = U diag(s) (guard inactive, since
    ||X^T X||_inf = max s^2 < 2) goes to U diag(s (15 - 10 s^2 + 3 s^4) / 8); iterated it reaches the
    polar factor from near-orthonormal and from badly scaled starts; from 1 + d it leaves a
    residual of order d^3 where one cubic step leaves order d^2."""
    rng = np.random.default_rng(8)
    k = 10
    U, _ = np.linalg.qr(rng.standard_normal((k, k)))
    s = np.linspace(0.5, 1.2, k)
    got = core.polar_renorm(U * s, 1, degree=5)
    np.testing.assert_allclose(got, U * (s * (15 - 10 * s ** 2 + 3 * s ** 4) / 8), rtol=0, atol=1e-14)
    near = U + 1e-3 * rng.standard_normal((k, k))
    far = U @ np.diag(np.geomspace(0.3, 80.0, k)) @ np.linalg.qr(rng.standard_normal((k, k)))[0]
    for X, iters in ((near, 3), (far, 30)):
        Uu, _, Vt = np.linalg.svd(X)
        assert np.abs(core.polar_renorm(X, iters, degree=5) - Uu @ Vt).max() < 1e-10
    D = rng.standard_normal((k, k))

    def residual(eps, degree):
        x = core.polar_renorm(U + eps * D, 1, degree=degree)
        return np.abs(x.T @ x - np.eye(k)).max()
    # halving the perturbation divides the residual by ~8 (degree 5) and ~4 (degree 3)
    assert 6.0 < residual(2e-3, 5) / residual(1e-3, 5) < 10.0
    assert 3.0 < residual(2e-3, 3) / residual(1e-3, 3) < 5.0
    assert residual(1e-3, 5) < 0.05 * residual(1e-3, 3)
    with pytest.raises(ValueError):
        core.polar_renorm(near, 1, degree=4)
