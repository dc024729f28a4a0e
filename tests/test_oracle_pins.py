"""Pins the oracle functions that every GPU parity test uses as its checker, each against an
independent statement of what the reference computes: dense formulas of the reference's own
buffers, loop-for-loop transcriptions of its code, closed forms and library factorisations. Each
check is chosen so that a plausible mistake in the oracle (swapped operand, wrong exponent, dropped
term) fails it."""
import numpy as np
import pytest

from oracle import core, gen


def _graph(seed=0, n=(260, 240, 250, 230), k=4, p_in=0.5, p_out=0.03):
    p = gen.GenParams(*n, k, 1.0, p_in, p_out, seed=seed)
    g = gen.generate(p)
    return {nm: core.Csr(rp, ci, nr, nc) for nm, (rp, ci, nr, nc) in g.blocks.items()}


@pytest.fixture(scope="module")
def whitened():
    # unequal partitions: a transposed or swapped operand changes a shape or a value
    bl = _graph()
    cfg = core.WhitenConfig(k=4, alpha0=1.0, oversample=10, power_iters=8, seed=1)
    wr = core.whiten(bl["XA"], bl["XB"], bl["XC"], cfg)
    we = core.whiten_exact(bl["XA"], bl["XB"], bl["XC"], cfg)
    return bl, wr, we


def test_factored_pseudo_inverse_equals_dense(whitened):
    """Z_B = L_B U^T and Z_C = L_C V^T (the factored forms) equal P_AC P_BC^+ and P_AB P_CB^+ formed
    densely from the exact SVD (pre_proc.c:215-218, 269)."""
    bl, wr, we = whitened
    ZB = wr["L_B"] @ wr["U"].T
    ZC = wr["L_C"] @ wr["V"].T
    assert np.abs(ZB - we["Z_B"]).max() < 1e-8 * np.abs(we["Z_B"]).max()
    assert np.abs(ZC - we["Z_C"]).max() < 1e-8 * np.abs(we["Z_C"]).max()
    # L_B pairs with U and L_C with V: the crossed products are far from Z_B, Z_C on the rows they share
    m = min(wr["U"].shape[0], wr["V"].shape[0])
    crossed = wr["L_B"] @ wr["V"][:m].T - we["Z_B"][:, :m]
    assert np.abs(crossed).max() > 1e-3 * np.abs(we["Z_B"]).max()


def test_yvectors_equal_dense_whitened_neighbours(whitened):
    """Y_A = G_XA W, Y_B = G_XB Z_B^T W, Y_C = G_XC Z_C^T W and the whitened means W^T mu_A,
    W^T Z_B mu_B, W^T Z_C mu_C, formed from the dense Z_B, Z_C of the exact whitening
    (pre_proc.c:339-366: Y_B = W^T Z_B G_XB^T is k x n_X, here node-major)."""
    bl, wr, we = whitened
    y = core.yvectors(bl["XA"], bl["XB"], bl["XC"], wr)
    W = we["W"]
    A, B, C = bl["XA"].dense(), bl["XB"].dense(), bl["XC"].dense()
    nx = A.shape[0]
    want = {"Y_A": A @ W, "Y_B": B @ we["Z_B"].T @ W, "Y_C": C @ we["Z_C"].T @ W,
            "mu_A": W.T @ A.mean(0), "mu_B": W.T @ we["Z_B"] @ B.mean(0), "mu_C": W.T @ we["Z_C"] @ C.mean(0)}
    for nm, v in want.items():
        assert np.abs(y[nm] - v).max() < 1e-7 * np.abs(v).max(), nm
    assert y["Y_A"].shape == (nx, 4)
    # n_B != n_C here, so exchanging the B and C factors (U for V) cannot even be formed
    assert wr["U"].shape[0] != wr["V"].shape[0]


def test_post_proc_equals_dense_reference_formulas(whitened):
    """Pi_X = L' Phi^T W^T G_XA^T, Pi_A = L' Phi^T W^T Z_B G_BA, Pi_B = L' Phi^T W^T G_BA^T,
    Pi_C = L' Phi^T W^T G_CA^T (post_proc.c:87-106; k x N there, node-major [N, k] here) with
    a non-orthonormal Phi and a non-uniform Lambda'."""
    bl, wr, we = whitened
    rng = np.random.default_rng(3)
    phi = rng.standard_normal((4, 4))
    lam, lam_p, _ = core.eigenvalues(phi)
    assert np.ptp(lam_p) > 0.1
    Pis = core.post_proc(wr, phi, lam_p, bl["XA"], bl["BA"], bl["CA"])
    U = we["W"] @ phi @ np.diag(lam_p)             # W Phi Lambda'  (post_proc.c:87-90)
    GXA, GBA, GCA = bl["XA"].dense(), bl["BA"].dense(), bl["CA"].dense()
    # G_BA is n_B x n_A (rows = B nodes) and Z_B is n_A x n_B, so Pi_A = U^T Z_B G_BA is k x n_A
    want = [(U.T @ GXA.T).T, (U.T @ we["Z_B"] @ GBA).T, (U.T @ GBA.T).T, (U.T @ GCA.T).T]
    for i, (got, w) in enumerate(zip(Pis, want)):
        assert got.shape == w.shape, i
        assert np.abs(got - w).max() < 1e-7 * np.abs(w).max(), i


def test_eigenvalues_on_scaled_columns():
    """Phi = Q diag(s): lambda_i = ||phi_i||^3 = s_i^3 (stpm.c:799,809, a dgemv of each column with
    itself raised to 1.5), Lambda' = ||lambda||_2 / lambda_i (genInvEvalMat, wrappers.cu:74-96),
    alpha_hat_i = lambda_i^-2 / sum lambda_j^-2 (stpm.c:848-870)."""
    rng = np.random.default_rng(7)
    k = 6
    Q, _ = np.linalg.qr(rng.standard_normal((k, k)))
    s = np.array([0.5, 0.8, 1.0, 1.3, 1.7, 2.5])
    lam, lam_p, ah = core.eigenvalues(Q * s)
    np.testing.assert_allclose(lam, s ** 3, rtol=1e-13)
    np.testing.assert_allclose(lam_p, np.linalg.norm(s ** 3) / s ** 3, rtol=1e-13)
    np.testing.assert_allclose(ah, s ** -6 / (s ** -6).sum(), rtol=1e-13)
    # the PINV_TOL cut of genInvEvalMat: a vanishing eigenvalue gets Lambda' = 0
    s2 = s.copy()
    s2[0] = 1e-3
    _, lam_p2, _ = core.eigenvalues(Q * s2)
    assert lam_p2[0] == 0.0 and lam_p2[1] > 0


def test_stgd_step_is_a_third_of_the_adjoint_on_symmetric_inputs():
    """With y_A = y_B = y_C and mu_A = mu_B = mu_C the three modes of T coincide, so the GPU update
    without renormalisation is Phi - eta * grad / 3 with grad the exact adjoint (itself pinned by
    finite differences), for a non-orthonormal Phi (the theta term is not Phi) and theta != 1."""
    rng = np.random.default_rng(11)
    k, b = 5, 9
    Y = rng.standard_normal((b, k))
    mu = rng.standard_normal(k)
    phi = rng.standard_normal((k, k))
    for scaling in ("gpu", "paper"):
        new, pre = core.stgd_step(phi, Y, Y, Y, mu, mu, mu, alpha0=1.7, eta=3e-3, theta=2.5, renorm=False,
                                  scaling=scaling)
        g = core.tensor_grad(phi, Y, Y, Y, mu, mu, mu, core.coeffs_for(1.7, scaling), 2.5)
        np.testing.assert_allclose(new, phi - 3e-3 * g / 3.0, rtol=1e-12, atol=1e-13)
        assert new is pre


def test_stgd_step_matches_reference_analytic_gradient():
    """The single-node GPU step (stpm.c:649-691, init_buff scaling s = 2/((a0+1)(a0+2))) equals
    phi - eta * s * Diff_Loss / 3 with theta_ref = theta / s, where Diff_Loss is the loop-for-loop
    transcription of the CPU analytic gradient (Util.cpp:257-304) on one symmetric sample."""
    rng = np.random.default_rng(12)
    k = 4
    y = rng.standard_normal(k)
    mu = rng.standard_normal(k)
    phi = rng.standard_normal((k, k))
    a0, eta, theta = 0.9, 1e-3, 1.0
    s = 2.0 / ((a0 + 1.0) * (a0 + 2.0))
    new, _ = core.stgd_step(phi, y[None], y[None], y[None], mu, mu, mu, a0, eta, theta, renorm=False, scaling="gpu")
    ref = core.ref_diff_loss_vectors(phi, y, y, y, mu, mu, mu, a0, theta=theta / s)
    np.testing.assert_allclose(new, phi - eta * s * ref / 3.0, rtol=1e-12, atol=1e-13)
    # the paper scaling is the same update with T multiplied by (a0+1)(a0+2)/2
    newp, _ = core.stgd_step(phi, y[None], y[None], y[None], mu, mu, mu, a0, eta, theta, renorm=False,
                             scaling="paper")
    refp = core.ref_diff_loss_vectors(phi, y, y, y, mu, mu, mu, a0, theta=theta)
    np.testing.assert_allclose(newp, phi - eta * refp / 3.0, rtol=1e-12, atol=1e-13)


def test_coefficients_pair_with_m2_scaling():
    a0 = 2.0
    assert core.coeffs_for(a0, "gpu") == core.gpu_coeffs(a0)
    assert core.coeffs_for(a0, "paper") == core.paper_coeffs(a0)
    # paper = (a0+1)(a0+2)/2 x gpu, term by term
    np.testing.assert_allclose(np.array(core.paper_coeffs(a0)) / np.array(core.gpu_coeffs(a0)),
                               (a0 + 1) * (a0 + 2) / 2)
    with pytest.raises(ValueError):
        core.coeffs_for(a0, "cpu")


def test_qr_renorm_is_householder_qr_with_positive_diagonal():
    """culaDeviceDgeqrfp + Dorgqr (stpm.c:774-786): Q with R = Q^T Phi upper triangular with a
    positive diagonal; for a near-orthonormal Phi it differs from the polar factor at first order."""
    rng = np.random.default_rng(13)
    k = 7
    phi = rng.standard_normal((k, k))
    Q = core.qr_renorm(phi)
    R = Q.T @ phi
    assert np.abs(Q.T @ Q - np.eye(k)).max() < 1e-13
    assert np.abs(np.tril(R, -1)).max() < 1e-12 and (np.diag(R) > 0).all()
    # Gram-Schmidt of the columns gives the same Q (uniqueness of the positive-diagonal QR)
    G = np.zeros_like(phi)
    for j in range(k):
        v = phi[:, j] - G[:, :j] @ (G[:, :j].T @ phi[:, j])
        v = v - G[:, :j] @ (G[:, :j].T @ v)
        G[:, j] = v / np.linalg.norm(v)
    np.testing.assert_allclose(Q, G, atol=1e-12)
    near = np.linalg.qr(rng.standard_normal((k, k)))[0] + 1e-3 * rng.standard_normal((k, k))
    assert np.abs(core.qr_renorm(near) - core.polar_renorm(near, 6)).max() > 1e-5


def test_stop_rule_follows_the_reference_loop():
    """stpm.c:747-765 with sgd_iter from 0 and incremented after the test: converged when
    frob < THRESH and sgd_iter > SGD_ITER_MIN, capped when sgd_iter > SGD_ITER_MAX; the loop ends
    after the step that meets either."""
    assert not core.stop_rule(5, 0.5, 1.0, 5, 0) and core.stop_rule(6, 0.5, 1.0, 5, 0)
    assert not core.stop_rule(6, 1.0, 1.0, 5, 0)            # not below THRESH
    assert not core.stop_rule(7, 9.0, 1.0, 5, 7) and core.stop_rule(8, 9.0, 1.0, 5, 7)
    # the reference defaults THRESH = 1, SGD_ITER_MIN = NX, SGD_ITER_MAX = NX + 2 (stpm.h:94-97) with
    # a frob that never drops below 1: the loop runs NX + 4 steps
    NX = 10
    steps = next(t for t in range(100) if core.stop_rule(t, 5.0, 1.0, NX, NX + 2)) + 1
    assert steps == NX + 4
    # a transcription of the loop itself agrees
    t, done = 0, False
    while not done:
        frob = 5.0
        done = (frob < 1.0 and t > NX) or (t > NX + 2)
        t += 1
    assert t == steps


def test_stgd_qr_mode_logs_the_change_before_the_qr_and_stops():
    rng = np.random.default_rng(14)
    k, n = 4, 40
    Y = [rng.standard_normal((n, k)) for _ in range(3)]
    mu = [rng.standard_normal(k) * 0.1 for _ in range(3)]
    phi0 = core.qr_renorm(rng.standard_normal((k, k)))
    phi, log = core.stgd(*Y, *mu, phi0, alpha0=1.0, eta=1e-2, batch=4, epochs=1, scaling="gpu", renorm_mode="qr")
    # replay by hand: frob is ||Phi~ - Phi||^2, then Phi <- Q(Phi~)
    p = phi0
    for s in range(10):
        sl = slice(4 * s, 4 * s + 4)
        c = core.gpu_coeffs(1.0)
        G = p.T @ p
        pre = p - 1e-2 * (p @ (G * G) - core.t_hat_c(p, Y[0][sl], Y[1][sl], Y[2][sl], *mu, c))
        assert np.isclose(log[s], ((pre - p) ** 2).sum(), rtol=1e-12)
        p = core.qr_renorm(pre)
    np.testing.assert_allclose(phi, p, atol=1e-12)
    # THRESH above every logged value with SGD_ITER_MIN = 3: ends after step index 4 (5 steps)
    _, log2 = core.stgd(*Y, *mu, phi0, alpha0=1.0, eta=1e-2, batch=4, epochs=1, scaling="gpu", renorm_mode="qr",
                        thresh=1e9, iter_min=3)
    assert log2.size == 5
    _, log3 = core.stgd(*Y, *mu, phi0, alpha0=1.0, eta=1e-2, batch=4, epochs=3, scaling="gpu", thresh=0.0,
                        iter_max=6)
    assert log3.size == 8


def _tensor_form_power_iteration(Da, Db, Dc, mua, mub, muc, phi, a0):
    """Util.cpp:153-166 transcribed: D_* are k x n_X (columns = samples), Pair_** = D_* D_*^T,
    para_main = (a0+1)(a0+2)/2/n_X, para_shift1 = -a0(a0+1)/2/n_X, para_shift0 = a0^2."""
    nx = Da.shape[1]
    pm, p1, p0 = (a0 + 1) * (a0 + 2) / 2 / nx, -a0 * (a0 + 1) / 2 / nx, a0 * a0
    Pab, Pac, Pbc = Da @ Db.T, Da @ Dc.T, Db @ Dc.T
    new = np.zeros_like(phi)
    for j in range(phi.shape[1]):
        v = phi[:, j]
        main = Da @ ((Db.T @ v) * (Dc.T @ v))                                      # tensor_form_main
        shift1 = (v @ Pbc @ v) * mua + (v @ mub) * (Pac @ v) + (v @ muc) * (Pab @ v)  # tensor_form_shift1
        shift0 = (v @ mub) * (v @ muc) * mua                                        # tensor_form_shift0
        new[:, j] = pm * main + p1 * shift1 + p0 * shift0
    lam = ((new ** 2).sum(0)) ** 1.5
    return new / np.sqrt((new ** 2).sum(0)), lam


def test_power_method_matches_reference_transcription():
    rng = np.random.default_rng(15)
    k, n = 4, 30
    YA, YB, YC = (rng.standard_normal((n, k)) for _ in range(3))
    mua, mub, muc = (rng.standard_normal(k) * 0.3 for _ in range(3))
    phi0 = core.qr_renorm(rng.standard_normal((k, k)))
    phi, lam = core.power_method(YA, YB, YC, mua, mub, muc, phi0, alpha0=0.8, iters=3, orthogonalize=False)
    p = phi0
    for _ in range(3):
        p, lam_ref = _tensor_form_power_iteration(YA.T, YB.T, YC.T, mua, mub, muc, p, 0.8)
    np.testing.assert_allclose(phi, p, rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(lam, lam_ref, rtol=1e-11)


def test_power_method_orthonormalisation_is_the_polar_factor():
    """The orthonormalisation is U V^T of Phi's SVD (the nearest orthonormal matrix), also when the
    columns are nearly dependent (cond ~ 1e7), where forming (Phi^T Phi)^-1/2 loses about half the
    digits."""
    rng = np.random.default_rng(16)
    k, n = 5, 50
    YA, YB, YC = (rng.standard_normal((n, k)) for _ in range(3))
    z = np.zeros(k)
    phi0 = core.qr_renorm(rng.standard_normal((k, k)))
    phi, _ = core.power_method(YA, YB, YC, z, z, z, phi0, alpha0=1.0, iters=1, orthogonalize=True)
    raw, _ = core.power_method(YA, YB, YC, z, z, z, phi0, alpha0=1.0, iters=1, orthogonalize=False)
    U, _, Vt = np.linalg.svd(raw)
    np.testing.assert_allclose(phi, U @ Vt, atol=1e-12)
    # polar factor = argmin ||Q - raw||_F over orthonormal Q: no nearby orthonormal matrix is closer
    for _ in range(5):
        Qp = core.qr_renorm(phi + 1e-3 * rng.standard_normal((k, k)))
        assert np.linalg.norm(phi - raw) <= np.linalg.norm(Qp - raw) + 1e-12
