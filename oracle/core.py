"""FP64 restatement of the MMSB tensor-method pipeline (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module, as the
checker or the timed CPU baseline. It follows the intended mathematics of the reference
(pre_proc.c:112-372, gemm_pinv_nys.c:207-421, stpm.c:497-870, post_proc.c:62-119,
CPU Util.cpp:38-304, Probability.cpp:33-99) with the factored, sparse formulation that the HIP
library implements: adjacency blocks stay in CSR, pair matrices P_** are applied as SpMM chains,
pseudo-inverses are rank-k randomized SVDs, and M2 is whitened through a randomized range finder
and a Rayleigh-Ritz step. Everything here is float64; Gaussian test matrices may be rounded to
float32 so that they equal the library's copies.

Conventions: tall matrices are row-major [n, cols] (one row per node); Phi is k x k with column i
the i-th eigenvector (the reference's column-major evec buffer, stpm.c:935-937).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .philox import STREAM_OMEGA, draw, u53, u53_open0


# ---- sparse helpers ------------------------------------------------------------------------
def csr_spmm(rowptr, colidx, vals, B, m, alpha=1.0):
    """Y = alpha * G @ B for CSR G with m rows (float64; each output row accumulates its nonzeros
    in CSR order, as scipy's CSR-times-dense loop does)."""
    import scipy.sparse as sp
    B = np.asarray(B, np.float64)
    rowptr = np.asarray(rowptr, np.int64)
    r0, r1 = int(rowptr[0]), int(rowptr[m])
    if r1 == r0:
        return np.zeros((m, B.shape[1]))
    cols = np.asarray(colidx, np.int64)[r0:r1]
    w = np.ones(r1 - r0) if vals is None else np.asarray(vals, np.float64)[r0:r1]
    G = sp.csr_matrix((w, cols, rowptr[: m + 1] - r0), shape=(m, B.shape[0]))
    return alpha * (G @ B)


def csr_transpose(rowptr, colidx, n_rows, n_cols, vals=None):
    rowptr = np.asarray(rowptr, np.int64)
    rows = np.repeat(np.arange(n_rows), np.diff(rowptr))
    cols = np.asarray(colidx, np.int64)
    order = np.lexsort((rows, cols))
    t_rowptr = np.zeros(n_cols + 1, np.int64)
    np.add.at(t_rowptr, cols + 1, 1)
    t_vals = None if vals is None else np.asarray(vals)[order]
    return np.cumsum(t_rowptr).astype(np.int32), rows[order].astype(np.int32), t_vals


def csr_dense(rowptr, colidx, n_rows, n_cols, vals=None):
    D = np.zeros((n_rows, n_cols))
    rows = np.repeat(np.arange(n_rows), np.diff(np.asarray(rowptr, np.int64)))
    D[rows, np.asarray(colidx, np.int64)] = 1.0 if vals is None else vals
    return D


@dataclass
class Csr:
    rowptr: np.ndarray
    colidx: np.ndarray
    n_rows: int
    n_cols: int
    vals: np.ndarray | None = None

    def mul(self, B, alpha=1.0):
        return csr_spmm(self.rowptr, self.colidx, self.vals, B, self.n_rows, alpha)

    def T(self) -> "Csr":
        rp, ci, v = csr_transpose(self.rowptr, self.colidx, self.n_rows, self.n_cols, self.vals)
        return Csr(rp, ci, self.n_cols, self.n_rows, v)

    def dense(self):
        return csr_dense(self.rowptr, self.colidx, self.n_rows, self.n_cols, self.vals)

    def colsum(self):
        return np.bincount(np.asarray(self.colidx, np.int64), weights=None if self.vals is None else self.vals,
                           minlength=self.n_cols).astype(np.float64)


# ---- random test matrices (replaces gen_rnd_mat, curand_rand_mat.cu:30-53) ------------------
def gaussian_matrix(seed: int, index: int, n: int, r: int, fp32: bool = True) -> np.ndarray:
    """N(0,1) matrix [n, r]: element (i, j) from Philox counter (i, j, STREAM_OMEGA+index, 0),
    Box-Muller z = sqrt(-2 ln u1) cos(2 pi u2), u1 in (0,1], u2 in [0,1)."""
    i = np.repeat(np.arange(n, dtype=np.uint32), r)
    j = np.tile(np.arange(r, dtype=np.uint32), n)
    w = draw(seed, STREAM_OMEGA + index, i, j, 0)
    u1 = u53_open0(w[0], w[1])
    u2 = u53(w[2], w[3])
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    z = z.reshape(n, r)
    return z.astype(np.float32).astype(np.float64) if fp32 else z


# ---- orthonormalisation (tall-skinny; replaces culaDgeqrfp/culaDorgqr, gemm_pinv_nys.c:284-286,385-394)
CHOL_SHIFT = 1e-14


def chol_upper(G: np.ndarray) -> np.ndarray:
    """Upper Cholesky factor of G + s I, s = CHOL_SHIFT * trace(G) (keeps rank-deficient Grams PD)."""
    s = CHOL_SHIFT * float(np.trace(G))
    if s <= 0.0:
        s = 1e-300
    L = np.linalg.cholesky(G + s * np.eye(G.shape[0]))
    return L.T


def cholqr2(X: np.ndarray):
    """Two passes of CholeskyQR: X = Q R with R = R2 R1 upper triangular."""
    R1 = chol_upper(X.T @ X)
    Q1 = X @ np.linalg.inv(R1)
    R2 = chol_upper(Q1.T @ Q1)
    Q = Q1 @ np.linalg.inv(R2)
    return Q, R2 @ R1


def col_sign_fix(M: np.ndarray):
    """Flip columns so that each column's largest-magnitude entry (first on ties) is positive."""
    idx = np.argmax(np.abs(M), axis=0)
    s = np.sign(M[idx, np.arange(M.shape[1])])
    s[s == 0] = 1.0
    return M * s, s


@dataclass
class WhitenConfig:
    k: int
    alpha0: float
    oversample: int = 10
    power_iters: int = 1
    seed: int = 0
    pinv_tol: float = 1e-6        # PINV_TOL, stpm.h:95 (relative to the largest value here)
    m2_scaling: str = "paper"      # "paper": CPU Util.cpp:81,109-113; "gpu": pre_proc.c:292-302
    fp32_omega: bool = True


# ---- whitening (pre_proc.c:112-372 / CPU second_whiten Util.cpp:38-127) ----------------------
def whiten(G_XA: Csr, G_XB: Csr, G_XC: Csr, cfg: WhitenConfig, progress=None) -> dict:
    """progress(stage), when given, is called after each stage (long runs report where they are)."""
    say = progress or (lambda _msg: None)
    nx = G_XA.n_rows
    k, r = cfg.k, cfg.k + cfg.oversample
    inv = 1.0 / nx
    mu_A = G_XA.colsum() * inv                         # pre_proc.c:154
    mu_B = G_XB.colsum() * inv                         # pre_proc.c:158
    mu_C = G_XC.colsum() * inv                         # pre_proc.c:162
    T_XA, T_XB, T_XC = G_XA.T(), G_XB.T(), G_XC.T()
    P_BC = lambda X: T_XB.mul(G_XC.mul(X), inv)        # (1/nx) G_XB^T G_XC X   (pre_proc.c:170)
    P_CB = lambda X: T_XC.mul(G_XB.mul(X), inv)

    # randomized SVD of P_BC (replaces pinv_nys_asym, gemm_pinv_nys.c:325-421)
    Om = gaussian_matrix(cfg.seed, 0, G_XC.n_cols, r, cfg.fp32_omega)
    Q, _ = cholqr2(P_BC(Om))
    say("whiten: P_BC sketch")
    for _ in range(cfg.power_iters):
        Qz, _ = cholqr2(P_CB(Q))
        Q, _ = cholqr2(P_BC(Qz))
        say("whiten: P_BC power iteration")
    Q2, R2 = cholqr2(P_CB(Q))                          # P_CB Q = Q2 R2  =>  P_BC ~ Q R2^T Q2^T
    say("whiten: P_BC factors")
    Ut, sig, Vth = np.linalg.svd(R2.T)
    U = Q @ Ut[:, :k]
    V = Q2 @ Vth.T[:, :k]
    sig = sig[:k]
    U, s = col_sign_fix(U)
    V = V * s
    sinv = np.where(sig > cfg.pinv_tol * sig[0], 1.0 / np.where(sig > 0, sig, 1.0), 0.0)

    # Z_B = P_AC P_BC^+ = L_B U^T, Z_C = P_AB P_CB^+ = L_C V^T   (pre_proc.c:215-218,269)
    L_B = T_XA.mul(G_XC.mul(V), inv) * sinv
    say("whiten: L_B")
    L_C = T_XA.mul(G_XB.mul(U), inv) * sinv
    say("whiten: L_C")

    # M2 as an operator on n_A-vectors (pre_proc.c:292-302; Util.cpp:81-113), symmetrised
    if cfg.m2_scaling == "paper":
        s_main, s_shift, diag_back = cfg.alpha0 + 1.0, cfg.alpha0, False
    else:
        s_main, s_shift, diag_back = 1.0, cfg.alpha0 / (cfg.alpha0 + 1.0), True

    def M2(X):
        out = 0.5 * s_main * (L_C @ (sig[:, None] * (L_B.T @ X)) + L_B @ (sig[:, None] * (L_C.T @ X)))
        out -= s_shift * np.outer(mu_A, mu_A @ X)
        if diag_back:
            out += s_shift * (mu_A * mu_A)[:, None] * X
        return out

    Om2 = gaussian_matrix(cfg.seed, 1, G_XA.n_cols, r, cfg.fp32_omega)
    Qm, _ = cholqr2(M2(Om2))
    for _ in range(cfg.power_iters):
        Qm, _ = cholqr2(M2(Qm))
    say("whiten: M2 range")
    S = Qm.T @ M2(Qm)
    S = 0.5 * (S + S.T)
    lam, E = np.linalg.eigh(S)
    order = np.argsort(-lam, kind="stable")[:k]
    lam, E = lam[order], E[:, order]
    lam_inv_sqrt = np.where(lam > cfg.pinv_tol * max(lam[0], 0.0), 1.0 / np.sqrt(np.where(lam > 0, lam, 1.0)), 0.0)
    W = (Qm @ E) * lam_inv_sqrt                        # W = U_k Lambda_k^-1/2  (Util.cpp:116-123)
    W, _ = col_sign_fix(W)
    return dict(W=W, mu_A=mu_A, mu_B=mu_B, mu_C=mu_C, U=U, V=V, sigma=sig, sigma_inv=sinv,
                L_B=L_B, L_C=L_C, lam_M2=lam, M2=M2)


def whiten_exact(G_XA: Csr, G_XB: Csr, G_XC: Csr, cfg: WhitenConfig) -> dict:
    """Dense FP64 whitening with exact SVD / eigendecomposition (small sizes only)."""
    nx = G_XA.n_rows
    k = cfg.k
    A, B, C = G_XA.dense(), G_XB.dense(), G_XC.dense()
    mu_A = A.mean(0)
    P_AC, P_BC, P_AB = A.T @ C / nx, B.T @ C / nx, A.T @ B / nx
    Uf, sf, Vfh = np.linalg.svd(P_BC)
    U, V, sig = Uf[:, :k], Vfh.T[:, :k], sf[:k]
    sinv = np.where(sig > cfg.pinv_tol * sig[0], 1.0 / sig, 0.0)
    Z_B = P_AC @ (V * sinv) @ U.T
    Z_C = P_AB @ (U * sinv) @ V.T
    P_CB = P_BC.T
    if cfg.m2_scaling == "paper":
        M2 = (cfg.alpha0 + 1.0) * Z_C @ P_CB @ Z_B.T - cfg.alpha0 * np.outer(mu_A, mu_A)
    else:
        mm = np.outer(mu_A, mu_A)
        np.fill_diagonal(mm, 0.0)
        M2 = Z_C @ P_CB @ Z_B.T - cfg.alpha0 / (cfg.alpha0 + 1.0) * mm
    M2 = 0.5 * (M2 + M2.T)
    lam, E = np.linalg.eigh(M2)
    order = np.argsort(-lam, kind="stable")[:k]
    lam, E = lam[order], E[:, order]
    W = E / np.sqrt(lam)
    W, _ = col_sign_fix(W)
    return dict(W=W, M2=M2, lam_M2=lam, Z_B=Z_B, Z_C=Z_C, U=U, V=V, sigma=sig, mu_A=mu_A)


def yvectors(G_XA: Csr, G_XB: Csr, G_XC: Csr, wh: dict) -> dict:
    """Whitened neighbour vectors (pre_proc.c:339-366, CPU CommunityDetection.cpp:66-68)."""
    W = wh["W"]
    M_B = wh["U"] @ (wh["L_B"].T @ W)          # Z_B^T W
    M_C = wh["V"] @ (wh["L_C"].T @ W)          # Z_C^T W
    return dict(Y_A=G_XA.mul(W), Y_B=G_XB.mul(M_B), Y_C=G_XC.mul(M_C),
                mu_A=W.T @ wh["mu_A"], mu_B=M_B.T @ wh["mu_B"], mu_C=M_C.T @ wh["mu_C"], M_B=M_B, M_C=M_C)


# ---- STGD (stpm.c:497-790) -----------------------------------------------------------------
def gpu_coeffs(alpha0: float):
    """(c3, c1, c0) of T in the GPU scaling s = 2/((a+1)(a+2)) (init_buff, stpm.c:1026-1038)."""
    return 1.0, -alpha0 / (alpha0 + 2.0), 2.0 * alpha0 * alpha0 / ((alpha0 + 1.0) * (alpha0 + 2.0))


def paper_coeffs(alpha0: float):
    """(c3, c1, c0) of T in the paper / CPU scaling (The_second_term, Util.cpp:289-298)."""
    return (alpha0 + 1.0) * (alpha0 + 2.0) / 2.0, -alpha0 * (alpha0 + 1.0) / 2.0, alpha0 * alpha0


def coeffs_for(alpha0: float, scaling: str = "paper"):
    """The tensor coefficients that pair with an M2 scaling: the paper / CPU M2 (Util.cpp:81,109-113)
    goes with the paper T (Util.cpp:289-298), the GPU M2 (pre_proc.c:292-302) with the init_buff T
    (stpm.c:1026-1038), as each reference path does."""
    if scaling == "paper":
        return paper_coeffs(alpha0)
    if scaling == "gpu":
        return gpu_coeffs(alpha0)
    raise ValueError(scaling)


def init_phi(seed: int, k: int) -> np.ndarray:
    """Phi0: orthonormalised k x k Gaussian (replaces the time-seeded QR of rand(), stpm.c:1041-1064)."""
    Q, _ = cholqr2(gaussian_matrix(seed, 2, k, k, True))
    return Q


def t_hat_c(phi, YA, YB, YC, muA, muB, muC, coeffs):
    """T(phi_i, phi_i, I) for all i over a minibatch: the matrix the reference builds as
    y_C c_g^T + mu_C c_mu^T (stpm.c:532-635), averaged over the batch rows."""
    c3, c1, c0 = coeffs
    a, b = YA @ phi, YB @ phi
    ma, mb = muA @ phi, muB @ phi
    cg = c3 * a * b + c1 * (a * mb + ma * b)
    cmu = c0 * ma * mb + c1 * (a * b).mean(0)
    return YC.T @ cg / YA.shape[0] + np.outer(muC, cmu)


def polar_renorm(phi, ns_iters=2, degree=3):
    """Orthonormal renormalisation by Newton-Schulz iterations towards the polar factor
    X (X^T X)^-1/2, with G = X^T X:
      degree 3: X <- c X (3I - c^2 G)/2                         (quadratic: 1 + d -> 1 - 1.5 d^2)
      degree 5: X <- c X (15I - 10 c^2 G + 3 c^4 G^2)/8          (cubic: 1 + d -> 1 + 2.5 d^3)
    Both map each singular value s of cX through a polynomial (s (3 - s^2)/2, s (15 - 10 s^2 +
    3 s^4)/8) whose fixed point 1 attracts (0, sqrt(3)) and (0, sqrt(7/3)) respectively. The scale
    c = min(1, sqrt(2/||G||_inf)) (infinity norm = max row sum of |G|, an upper bound of
    sigma_max(X)^2) keeps every singular value of cX at or below sqrt(2), inside both regions; the
    polar factor is scale-invariant, so c changes nothing when it is 1 and only rescues iterates
    that would otherwise diverge. The reference re-orthonormalises every step with a QR
    factorisation instead (stpm.c:774-786)."""
    if degree not in (3, 5):
        raise ValueError("Newton-Schulz degree must be 3 or 5")
    I = np.eye(phi.shape[1])
    for _ in range(ns_iters):
        G = phi.T @ phi
        ginf = np.abs(G).sum(1).max()
        c = np.sqrt(2.0 / ginf) if ginf > 2.0 else 1.0
        if degree == 3:
            phi = c * phi @ (1.5 * I - 0.5 * c * c * G)
        else:
            Q = -1.25 * c * c * G + 0.375 * c ** 4 * (G @ G)
            phi = c * phi @ (1.875 * I + Q)
    return phi


def qr_renorm(phi):
    """Q of the QR factorisation with a non-negative diagonal of R (culaDeviceDgeqrfp followed by
    culaDeviceDorgqr, stpm.c:774-786): numpy's Householder QR with the signs of Q's columns set so
    that R's diagonal is positive (the factorisation is then unique for a full-rank Phi)."""
    Q, R = np.linalg.qr(phi)
    d = np.sign(np.diag(R))
    d[d == 0] = 1.0
    return Q * d


def stgd_step(phi, YA, YB, YC, muA, muB, muC, alpha0, eta, theta=1.0, renorm=True, ns_iters=2, ns_degree=3,
              scaling="paper", renorm_mode="ns"):
    """Phi <- Phi - eta [theta Phi (Phi^T Phi)^o2 - T(Phi, Phi, I)] (stpm.c:649-691) with the
    coefficients of `scaling`, then the orthonormal renormalisation: Newton-Schulz towards the polar
    factor (renorm_mode "ns") or the QR of the reference (renorm_mode "qr", stpm.c:774-786).
    Returns (Phi_next, Phi~) with Phi~ the iterate before the renormalisation."""
    G = phi.T @ phi
    T = t_hat_c(phi, YA, YB, YC, muA, muB, muC, coeffs_for(alpha0, scaling))
    new = phi - eta * (theta * phi @ (G * G) - T)
    pre = new
    if renorm:
        if renorm_mode == "qr":
            new = qr_renorm(new)
        elif renorm_mode == "ns":
            new = polar_renorm(new, ns_iters, ns_degree)
        else:
            raise ValueError(renorm_mode)
    return new, pre


def stop_rule(step, frob, thresh, iter_min, iter_max):
    """The STGD loop's end test after step `step` (0-based, sgd_iter of stpm.c:194,769) with
    ||dPhi||_F^2 = frob: converged when frob < THRESH and step > SGD_ITER_MIN (stpm.c:747), or the
    cap when step > SGD_ITER_MAX (stpm.c:761). A bound <= 0 switches its rule off."""
    if iter_min > 0 and frob < thresh and step > iter_min:
        return True
    return iter_max > 0 and step > iter_max


def stgd(Y_A, Y_B, Y_C, mu_A, mu_B, mu_C, phi0, alpha0, eta, batch, epochs, theta=1.0,
         renorm_every=1, steps=None, ns_iters=2, ns_degree=3, scaling="paper", renorm_mode="ns",
         thresh=None, iter_min=0, iter_max=0):
    """Sequential minibatches over X (stpm.c:499-502,768); returns Phi and ||dPhi||_F^2 per step.
    The log holds ||Phi_{t+1} - Phi_t||^2 after the renormalisation, except in renorm_mode "qr",
    which logs the change before the QR as the reference does (stpm.c:705-743, QR at :774-786).
    With thresh given, the loop ends early by stop_rule (checked after every step)."""
    phi = phi0.copy()
    n = Y_A.shape[0]
    log = []
    step = 0
    total = steps if steps is not None else epochs * ((n + batch - 1) // batch)
    for s in range(total):
        lo = (s % ((n + batch - 1) // batch)) * batch
        hi = min(lo + batch, n)
        renorm = renorm_every > 0 and ((step + 1) % renorm_every == 0)
        new, pre = stgd_step(phi, Y_A[lo:hi], Y_B[lo:hi], Y_C[lo:hi], mu_A, mu_B, mu_C, alpha0, eta, theta, renorm,
                             ns_iters, ns_degree, scaling, renorm_mode)
        ref = pre if renorm_mode == "qr" else new
        log.append(float(((ref - phi) ** 2).sum()))
        phi = new
        if thresh is not None and stop_rule(step, log[-1], thresh, iter_min, iter_max):
            break
        step += 1
    return phi, np.array(log)


def eigenvalues(phi, pinv_tol=1e-6):
    """lambda_i = ||phi_i||^3 (stpm.c:799,809); Lambda' = diag(||lambda||_2 / lambda_i) with the
    PINV_TOL cut (genInvEvalMat, wrappers.cu:74-96); alpha_hat_i = lambda_i^-2 / sum_j lambda_j^-2
    (stpm.c:848-870)."""
    lam = (phi * phi).sum(0) ** 1.5
    nrm = np.sqrt((lam * lam).sum())
    lam_p = np.where(np.abs(lam) > pinv_tol, nrm / lam, 0.0)
    inv2 = 1.0 / (lam * lam)
    alpha_hat = inv2 / inv2.sum()
    return lam, lam_p, alpha_hat


# ---- post_proc (post_proc.c:62-119, CPU CommunityDetection.cpp:98-140) ------------------------
def post_proc(wh: dict, phi, lam_p, G_XA: Csr, G_BA: Csr, G_CA: Csr):
    """Pi_X, Pi_A, Pi_B, Pi_C as [n, k] (row = node), i.e. the transposes of the reference's
    k x N buffers, which are written node-major."""
    Upost = wh["W"] @ (phi * lam_p)                      # W Phi Lambda'   (post_proc.c:87-90)
    Pi_X = G_XA.mul(Upost)                                # (post_proc.c:93)
    ZBt_U = wh["U"] @ (wh["L_B"].T @ Upost)               # Z_B^T Upost     (post_proc.c:97)
    Pi_A = G_BA.T().mul(ZBt_U)                             # (post_proc.c:100)
    Pi_B = G_BA.mul(Upost)                                 # (post_proc.c:103)
    Pi_C = G_CA.mul(Upost)                                 # (post_proc.c:106; G_CA, not G_BA)
    return Pi_X, Pi_A, Pi_B, Pi_C


def norm_prob_rows(P: np.ndarray) -> np.ndarray:
    """normProbVector (Probability.cpp:33-57) applied to each node's k-vector (row here, column
    there): all-zero stays; pos = sum of v > 0, neg = sum of v <= 0; if |pos| < |neg| then
    v <- -v/|neg| else v <- v/|pos|; negatives clamped to 0."""
    P = np.asarray(P, np.float64)
    pos = np.where(P > 0, P, 0.0).sum(1)
    neg = np.where(P > 0, 0.0, P).sum(1)
    flip = np.abs(pos) < np.abs(neg)
    den = np.where(flip, np.abs(neg), np.abs(pos))
    zero = ~(P != 0).any(1)
    den = np.where(zero | (den == 0), 1.0, den)
    out = np.where(flip[:, None], -P, P) / den[:, None]
    out = np.where(out < 0, 0.0, out)
    out[zero] = 0.0
    return out


def memberships(Pn: np.ndarray, tau: float = 0.0):
    """Membership sets {c : Pi_hat[node, c] > tau} as a CSR (rowptr, community ids)."""
    mask = Pn > tau
    rowptr = np.r_[0, np.cumsum(mask.sum(1))].astype(np.int64)
    return rowptr, np.nonzero(mask)[1].astype(np.int32)


# ---- loss and adjoint (Appendix: exact gradient; reference analytic CPU Util.cpp:257-304) ------
def tensor_loss(phi, YA, YB, YC, muA, muB, muC, coeffs, theta):
    """L = theta/2 sum_ij <phi_i,phi_j>^3 - sum_i T_b(phi_i, phi_i, phi_i)."""
    c3, c1, c0 = coeffs
    a, b, c = YA @ phi, YB @ phi, YC @ phi
    ma, mb, mc = muA @ phi, muB @ phi, muC @ phi
    G = phi.T @ phi
    t = (c3 * a * b * c + c1 * (a * b * mc + a * mb * c + ma * b * c)).mean(0) + c0 * ma * mb * mc
    return 0.5 * theta * (G ** 3).sum() - t.sum()


def tensor_grad(phi, YA, YB, YC, muA, muB, muC, coeffs, theta):
    """dL/dPhi = 3 theta Phi (G o G) - [T(I,phi,phi) + T(phi,I,phi) + T(phi,phi,I)] column-wise."""
    c3, c1, c0 = coeffs
    nb = YA.shape[0]
    a, b, c = YA @ phi, YB @ phi, YC @ phi
    ma, mb, mc = muA @ phi, muB @ phi, muC @ phi
    G = phi.T @ phi
    KA = c3 * b * c + c1 * (b * mc + mb * c)
    KB = c3 * a * c + c1 * (a * mc + ma * c)
    KC = c3 * a * b + c1 * (a * mb + ma * b)
    kA = c0 * mb * mc + c1 * (b * c).mean(0)
    kB = c0 * ma * mc + c1 * (a * c).mean(0)
    kC = c0 * ma * mb + c1 * (a * b).mean(0)
    T = (YA.T @ KA + YB.T @ KB + YC.T @ KC) / nb + np.outer(muA, kA) + np.outer(muB, kB) + np.outer(muC, kC)
    return 3.0 * theta * phi @ (G * G) - T


def ref_diff_loss_vectors(phi, ya, yb, yc, muA, muB, muC, alpha0, theta=1e4):
    """The reference analytic STGD gradient for one sample (Diff_Loss + The_second_term,
    CPU Util.cpp:257-304), restated loop-for-loop: column i is
    3 theta sum_j <phi_i,phi_j>^2 phi_j - 3 T_paper(phi_i, phi_i, I)."""
    k = phi.shape[1]
    para0 = (alpha0 + 1.0) * (alpha0 + 2.0) / 2.0
    para1 = alpha0 * alpha0
    para2 = -alpha0 * (alpha0 + 1.0) / 2.0
    out = np.zeros_like(phi)
    for i in range(k):
        v = phi[:, i]
        sq = (v @ phi) ** 2
        term1 = 3.0 * theta * (phi * sq).sum(1)
        second = (para0 * (v @ ya) * (v @ yb) * yc + para1 * (v @ muA) * (v @ muB) * muC
                  + para2 * (v @ ya) * (v @ yb) * muC + para2 * (v @ ya) * (v @ muB) * yc
                  + para2 * (v @ muA) * (v @ yb) * yc)
        out[:, i] = term1 - 3.0 * second
    return out


# ---- batch implicit power method (CPU tensorDecom_alpha0, Util.cpp:131-206) -------------------
def power_method(Y_A, Y_B, Y_C, mu_A, mu_B, mu_C, phi0, alpha0, iters, orthogonalize=True):
    """phi_i <- T(I, phi_i, phi_i) over all of X with the paper coefficients (A-mode free, as
    tensor_form_main/shift1/shift0 Util.cpp:187-206), then column normalisation (Util.cpp:166),
    plus a symmetric orthonormalisation when requested (the reference has none, so its columns
    collapse). Returns Phi and lambda = ||T(I,phi,phi)||^3 per column, the reference's
    (sum phi^2)^(3/2) of the unnormalised iterate (Util.cpp:164-165). The orthonormalisation is the
    polar factor U V^T of Phi's own SVD, so it does not square Phi's condition number."""
    c3, c1, c0 = paper_coeffs(alpha0)
    phi = phi0.copy()
    n = Y_A.shape[0]
    lam = np.zeros(phi.shape[1])
    for _ in range(iters):
        b, c = Y_B @ phi, Y_C @ phi
        mb, mc = mu_B @ phi, mu_C @ phi
        KA = c3 * b * c + c1 * (b * mc + mb * c)
        kA = c0 * mb * mc + c1 * (b * c).mean(0)
        new = Y_A.T @ KA / n + np.outer(mu_A, kA)
        nrm = np.sqrt((new * new).sum(0))
        lam = nrm ** 3
        new = new / np.where(nrm > 0, nrm, 1.0)
        if orthogonalize:
            U, _, Vt = np.linalg.svd(new)
            new = U @ Vt
        phi = new
    return phi, lam
