/* FP64 C / OpenMP restatement of the STGD loop of the MMSB tensor method -- TEST INFRASTRUCTURE ONLY.
 *
 * A second CPU restatement beside the numpy one in core.py (stgd, stgd_step, t_hat_c,
 * polar_renorm), written independently of it: the tests check the two against each other, and the
 * benchmark times this one, phase by phase, as the compiled CPU baseline. Nothing in the library
 * links or calls it.
 *
 * The loop follows CommunityDetectionSoftware_GPU/stpm.c: sequential minibatches over X
 * (stpm.c:499-502,768), T(phi_i, phi_i, I) over a minibatch as Y_C^T c_g / m + mu_C c_mu^T
 * (stpm.c:532-635), the step Phi <- Phi - eta [theta Phi (Phi^T Phi)^o2 - T] (stpm.c:649-691) and
 * ||Phi_{t+1} - Phi_t||_F^2 per step (stpm.c:705-743). The renormalisation after each step is the
 * guarded Newton-Schulz iteration towards the polar factor (degree 3 or 5, scale
 * c = min(1, sqrt(2 / ||G||_inf)), G = X^T X) in place of the reference's QR (stpm.c:774-786).
 *
 * Storage is row-major FP64 throughout: Y_* are n x k (one row per node), Phi is k x k.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { MR = 6, NR = 8 };  /* register tile: 6 x 8 doubles = 12 four-wide accumulators */
typedef double v4d __attribute__((vector_size(32)));
typedef double v4du __attribute__((vector_size(32), aligned(8)));   /* unaligned loads and stores */

enum { IB = 16, KC = 128 };  /* a task: IB row tiles x one column panel; q in blocks of KC */

/* C = A B with A(i, q) = A[i * sai + q * sap] (so one routine serves A and A^T), B and C
 * row-major. A task owns IB row tiles of one 8-column panel and walks q in blocks of KC, so the
 * panel's KC x 8 block of B stays in L1 across the task's row tiles and consecutive tasks of a
 * thread reuse the same rows of A. Each output element is one FMA chain in q order (the partial
 * sums pass through C exactly): the result depends neither on the thread count nor on the blocking. */
static void gemm(int64_t m, int64_t n, int64_t p, const double *A, int64_t sai, int64_t sap,
                 const double *B, int64_t ldb, double *C, int64_t ldc)
{
    const int64_t mt = (m + MR - 1) / MR, nt = (n + NR - 1) / NR;
    int64_t ib = IB;                 /* fewer row tiles per task when there are few tasks per thread */
    while (ib > 1 && ((mt + ib - 1) / ib) * nt < 4 * (int64_t)omp_get_max_threads())
        ib >>= 1;
    const int64_t mb = (mt + ib - 1) / ib;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t bi = 0; bi < mb; bi++) {
        for (int64_t tj = 0; tj < nt; tj++) {
            const int64_t j0 = tj * NR, nj = n - j0 < NR ? n - j0 : NR;
            const int64_t ti_end = (bi + 1) * ib < mt ? (bi + 1) * ib : mt;
            for (int64_t q0 = 0; q0 < p || q0 == 0; q0 += KC) {
                const int64_t q1 = q0 + KC < p ? q0 + KC : p;
                for (int64_t ti = bi * ib; ti < ti_end; ti++) {
                    const int64_t i0 = ti * MR, mi = m - i0 < MR ? m - i0 : MR;
                    if (mi == MR && nj == NR) {
                        v4d acc[MR][2];
                        for (int r = 0; r < MR; r++) {
                            if (q0 == 0) {
                                acc[r][0] = acc[r][1] = (v4d){0.0, 0.0, 0.0, 0.0};
                            } else {
                                acc[r][0] = *(const v4du *)(C + (i0 + r) * ldc + j0);
                                acc[r][1] = *(const v4du *)(C + (i0 + r) * ldc + j0 + 4);
                            }
                        }
                        const double *a0 = A + i0 * sai;
                        for (int64_t q = q0; q < q1; q++) {
                            const double *br = B + q * ldb + j0;
                            const v4d b0 = *(const v4du *)br, b1 = *(const v4du *)(br + 4);
                            const double *aq = a0 + q * sap;
                            for (int r = 0; r < MR; r++) {
                                const double av = aq[r * sai];
                                acc[r][0] += av * b0;
                                acc[r][1] += av * b1;
                            }
                        }
                        for (int r = 0; r < MR; r++) {
                            *(v4du *)(C + (i0 + r) * ldc + j0) = acc[r][0];
                            *(v4du *)(C + (i0 + r) * ldc + j0 + 4) = acc[r][1];
                        }
                    } else {          /* an edge tile: the same FMA chains on zero-padded operands */
                        double cb[MR][NR];
                        for (int r = 0; r < MR; r++)
                            for (int c = 0; c < NR; c++)
                                cb[r][c] = (q0 > 0 && r < mi && c < nj) ? C[(i0 + r) * ldc + j0 + c] : 0.0;
                        v4d acc[MR][2];
                        for (int r = 0; r < MR; r++) {
                            acc[r][0] = *(const v4du *)&cb[r][0];
                            acc[r][1] = *(const v4du *)&cb[r][4];
                        }
                        const double *a0 = A + i0 * sai;
                        for (int64_t q = q0; q < q1; q++) {
                            const double *br = B + q * ldb + j0;
                            double bt[NR] = {0.0};
                            for (int64_t c = 0; c < nj; c++)
                                bt[c] = br[c];
                            const v4d b0 = *(const v4du *)bt, b1 = *(const v4du *)(bt + 4);
                            const double *aq = a0 + q * sap;
                            for (int r = 0; r < MR; r++) {
                                const double av = r < mi ? aq[r * sai] : 0.0;
                                acc[r][0] += av * b0;
                                acc[r][1] += av * b1;
                            }
                        }
                        for (int r = 0; r < MR; r++) {
                            *(v4du *)&cb[r][0] = acc[r][0];
                            *(v4du *)&cb[r][4] = acc[r][1];
                        }
                        for (int64_t r = 0; r < mi; r++)
                            for (int64_t c = 0; c < nj; c++)
                                C[(i0 + r) * ldc + j0 + c] = cb[r][c];
                    }
                }
                if (q1 >= p)
                    break;
            }
        }
    }
}

/* C (n x n') = A^T B for row-major A (p x n) and B (p x n'): A^T is first laid out row-major in
 * at (n x p), so the product runs with unit-stride rows of A^T; the FMA chains are those of gemm. */
static void gemm_tn(int64_t n, int64_t n2, int64_t p, const double *A, int64_t lda, const double *B, int64_t ldb,
                    double *C, int64_t ldc, double *at)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        for (int64_t q = 0; q < p; q++)
            at[i * p + q] = A[q * lda + i];
    gemm(n, n2, p, at, p, 1, B, ldb, C, ldc);
}

/* y = x^T M for a length-k vector x and a k x k row-major M. */
static void vecmat(int32_t k, const double *x, const double *M, double *y)
{
#pragma omp parallel for schedule(static)
    for (int32_t j = 0; j < k; j++) {
        double s = 0.0;
        for (int32_t q = 0; q < k; q++)
            s += x[q] * M[(int64_t)q * k + j];
        y[j] = s;
    }
}

/* X <- polar-factor step(s) of X (k x k) by guarded Newton-Schulz iterations. */
static void renorm_ns(int32_t k, double *X, int32_t iters, int32_t degree, double *G, double *G2, double *M,
                      double *tmp, double *at)
{
    const int64_t kk = (int64_t)k * k;
    for (int32_t it = 0; it < iters; it++) {
        gemm_tn(k, k, k, X, k, X, k, G, k, at);              /* G = X^T X */
        double ginf = 0.0;
        for (int32_t i = 0; i < k; i++) {                     /* ||G||_inf: largest absolute row sum */
            double s = 0.0;
            for (int32_t j = 0; j < k; j++)
                s += fabs(G[(int64_t)i * k + j]);
            if (s > ginf)
                ginf = s;
        }
        const double c = ginf > 2.0 ? sqrt(2.0 / ginf) : 1.0, c2 = c * c;
        if (degree == 3) {
            /* M = (3 I - c^2 G) / 2 */
            for (int64_t e = 0; e < kk; e++)
                M[e] = -0.5 * c2 * G[e];
        } else {
            /* M = (15 I - 10 c^2 G + 3 c^4 G^2) / 8 */
            gemm(k, k, k, G, k, 1, G, k, G2, k);
            for (int64_t e = 0; e < kk; e++)
                M[e] = -1.25 * c2 * G[e] + 0.375 * c2 * c2 * G2[e];
        }
        const double d = degree == 3 ? 1.5 : 1.875;
        for (int32_t i = 0; i < k; i++)
            M[(int64_t)i * k + i] += d;
        gemm(k, k, k, X, k, 1, M, k, tmp, k);                 /* X <- c X M */
        for (int64_t e = 0; e < kk; e++)
            X[e] = c * tmp[e];
    }
}

/* `steps` STGD steps from phi (k x k, in/out). c3, c1, c0 are the coefficients of
 * T = c3 E[y_A (x) y_B (x) y_C] + c1 (mean-shift terms) + c0 mu_A (x) mu_B (x) mu_C, as t_hat_c in
 * core.py states them. log_out[s] receives ||Phi_{s+1} - Phi_s||_F^2 after the renormalisation
 * (may be NULL). phase_s[0..3], when given, accumulate wall seconds of: the contractions Y_A Phi
 * and Y_B Phi; the c_g / c_mu epilogue and Y_C^T c_g; the step (Phi^T Phi, Phi (Phi^T Phi)^o2,
 * the update); the renormalisation. threads > 0 sets the OpenMP thread count.
 * Returns 0, -1 on bad arguments, -2 when memory runs out. */
int stgd_cpu_f64(const double *ya, const double *yb, const double *yc, int64_t n, int32_t k,
                 const double *mu_a, const double *mu_b, const double *mu_c, double *phi,
                 double c3, double c1, double c0, double eta, double theta, int64_t batch, int64_t steps,
                 int32_t ns_iters, int32_t ns_degree, int32_t threads, double *log_out, double *phase_s)
{
    if (!ya || !yb || !yc || !mu_a || !mu_b || !mu_c || !phi || n <= 0 || k <= 0 || batch <= 0 || steps < 0 ||
        ns_iters < 0 || (ns_degree != 3 && ns_degree != 5))
        return -1;
    if (threads > 0)
        omp_set_num_threads(threads);
    const int64_t kk = (int64_t)k * k;
    const int64_t mb = batch < n ? batch : n;
    const int64_t nbatches = (n + batch - 1) / batch;
    const int64_t mk = (mb > k ? mb : k) * k;
    double *buf = malloc(sizeof(double) * (size_t)(3 * mb * k + 7 * kk + 4 * (int64_t)k + mk));
    if (!buf)
        return -2;
    double *A = buf, *B = A + mb * k, *CG = B + mb * k;
    double *G = CG + mb * k, *H = G + kk, *T = H + kk, *NEW = T + kk, *G2 = NEW + kk, *M = G2 + kk, *TMP = M + kk;
    double *ma = TMP + kk, *mbv = ma + k, *mab = mbv + k, *cmu = mab + k, *AT = cmu + k;
    double ph[4] = {0, 0, 0, 0};

    for (int64_t s = 0; s < steps; s++) {
        const int64_t lo = (s % nbatches) * batch;
        const int64_t hi = lo + batch < n ? lo + batch : n;
        const int64_t m = hi - lo;
        const double inv_m = 1.0 / (double)m;

        double t0 = omp_get_wtime();
        gemm(m, k, k, ya + lo * k, k, 1, phi, k, A, k);      /* a = Y_A[lo:hi] Phi */
        gemm(m, k, k, yb + lo * k, k, 1, phi, k, B, k);      /* b = Y_B[lo:hi] Phi */
        double t1 = omp_get_wtime();
        ph[0] += t1 - t0;

        vecmat(k, mu_a, phi, ma);                            /* mu_A^T Phi */
        vecmat(k, mu_b, phi, mbv);                           /* mu_B^T Phi */
#pragma omp parallel for schedule(static)
        for (int32_t j0 = 0; j0 < k; j0 += NR) {             /* column means of a o b, row order */
            const int32_t j1 = j0 + NR < k ? j0 + NR : k;
            double s2[NR] = {0.0};
            for (int64_t r = 0; r < m; r++)
                for (int32_t j = j0; j < j1; j++)
                    s2[j - j0] += A[r * k + j] * B[r * k + j];
            for (int32_t j = j0; j < j1; j++) {
                mab[j] = s2[j - j0] * inv_m;
                cmu[j] = c0 * ma[j] * mbv[j] + c1 * mab[j];
            }
        }
#pragma omp parallel for schedule(static)
        for (int64_t r = 0; r < m; r++)                      /* c_g = c3 a o b + c1 (a o mu_B Phi + mu_A Phi o b) */
            for (int32_t j = 0; j < k; j++) {
                const double a = A[r * k + j], b = B[r * k + j];
                CG[r * k + j] = c3 * a * b + c1 * (a * mbv[j] + ma[j] * b);
            }
        gemm_tn(k, k, m, yc + lo * k, k, CG, k, T, k, AT);   /* Y_C[lo:hi]^T c_g */
        for (int32_t i = 0; i < k; i++)
            for (int32_t j = 0; j < k; j++)
                T[(int64_t)i * k + j] = T[(int64_t)i * k + j] * inv_m + mu_c[i] * cmu[j];
        double t2 = omp_get_wtime();
        ph[1] += t2 - t1;

        gemm_tn(k, k, k, phi, k, phi, k, G, k, AT);          /* G = Phi^T Phi */
        for (int64_t e = 0; e < kk; e++)
            H[e] = G[e] * G[e];
        gemm(k, k, k, phi, k, 1, H, k, TMP, k);              /* Phi (Phi^T Phi)^o2 */
        for (int64_t e = 0; e < kk; e++)
            NEW[e] = phi[e] - eta * (theta * TMP[e] - T[e]);
        double t3 = omp_get_wtime();
        ph[2] += t3 - t2;

        renorm_ns(k, NEW, ns_iters, ns_degree, G, G2, M, TMP, AT);
        double d2 = 0.0;
        for (int64_t e = 0; e < kk; e++) {
            const double d = NEW[e] - phi[e];
            d2 += d * d;
            phi[e] = NEW[e];
        }
        if (log_out)
            log_out[s] = d2;
        ph[3] += omp_get_wtime() - t3;
    }
    if (phase_s)
        for (int i = 0; i < 4; i++)
            phase_s[i] += ph[i];
    free(buf);
    return 0;
}
