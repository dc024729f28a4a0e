/*
 * mmsb_stpm.h -- C-ABI of the MI355X-native MMSB tensor-method community-detection library.
 *
 * Drop-in boundary for the reference's GPU path (CommunityDetectionSoftware_GPU):
 *   - the compile-time configuration of stpm.h:47-106 becomes the runtime struct stpm_config,
 *     whose field names mirror the macros (NX..NCOM, ALPHA_0, LEARN_RATE, THRESH, PINV_TOL,
 *     SGD_ITER_MIN/MAX, BINARY/EXPECTED, SYNTHETIC, FOLDER_NAME);
 *   - the program-level entry points pre_proc() (pre_proc.c:112), GPU_Task() (stpm.c:132) and
 *     post_proc() (post_proc.c:62) keep their names and read/write the same files;
 *   - the per-kernel launchers of wrappers.cu:148-198 and the CULA/cuRAND call sites
 *     (stpm.c:532-809, gemm_pinv_nys.c:27-421, curand_rand_mat.cu:30-53) are replaced by the
 *     mmsb_* entry points below, which take explicit device pointers and a hipStream_t.
 *
 * Conventions: every function returns an int status (MMSB_OK == 0, negative on error) and never
 * calls exit(); mmsb_last_error() returns the message of the last failure on the calling thread.
 * Device pointers are plain pointers into HBM; `stream` is a hipStream_t passed as void* (NULL is
 * the default stream). Dense tall matrices are row-major (one row per node) with a leading
 * dimension `ld*` in elements; k x k matrices are stated per function.
 */
#ifndef MMSB_STPM_H
#define MMSB_STPM_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (replace exit(1) I/O errors and exit(2) CULA errors, pre_proc.c:54-58,96-109) */
enum {
  MMSB_OK = 0,
  MMSB_ERR_ARG = -1,      /* invalid argument / shape                                   */
  MMSB_ERR_IO = -2,       /* file missing or unreadable (reference: exit(1))            */
  MMSB_ERR_HIP = -3,      /* HIP runtime failure (reference: cula_exception -> exit(2)) */
  MMSB_ERR_NOMEM = -4,    /* host or device allocation failed                           */
  MMSB_ERR_NUMERIC = -5,  /* factorisation failed (not positive definite, no convergence)*/
  MMSB_ERR_STATE = -6     /* entry point called out of order                            */
};

/* Message of the last failing call on this thread ("" if none). */
const char* mmsb_last_error(void);
/* Library version string. */
const char* mmsb_version(void);

/* ------------------------------------------------------------------------------------------
 * Kernel-level entry points
 * ------------------------------------------------------------------------------------------ */

/*
 * CSR SpMM, FP32:  Y[m x r] = alpha * G[m x n] * B[n x r].
 * Replaces every sparse-times-dense product of the whitening, y-vector and Pi stages, which the
 * reference runs as dense n x n DGEMMs through culaBlockDgemm (gemm_pinv_nys.c:27-145; call
 * sites pre_proc.c:154-366, post_proc.c:87-106).
 * rowptr[m+1], colidx[nnz] (int32, 0-based, every colidx < n), vals[nnz] or NULL for a binary
 * graph (all ones). B and Y row-major with leading dimensions ldb, ldy (multiples of 4 floats,
 * 16-byte aligned base pointers), r <= ldb, r <= ldy. Y is overwritten.
 */
int mmsb_spmm_f32(void* stream, int32_t m, int32_t n, int64_t nnz, const int32_t* rowptr,
                  const int32_t* colidx, const float* vals, float alpha, const float* B,
                  int32_t ldb, int32_t r, float* Y, int32_t ldy);
/* The same product with the CSR block in host memory (rowptr, colidx, vals host pointers; B, Y on
 * the device): rows are cut into chunks of at most chunk_nnz nonzeros, staged through pinned memory
 * and copied on a copy stream while the previous chunk is multiplied on `stream` -- the blocking
 * idea of culaBlockDgemm (gemm_pinv_nys.c:27-145) applied to the sparse operand, for blocks not
 * kept in HBM. Y is identical to mmsb_spmm_f32's. The CSR is validated on the host (monotone
 * rowptr, column indices below n). Returns when Y is complete; calls are serialised. */
int mmsb_spmm_f32_host_csr(void* stream, int32_t m, int32_t n, int64_t nnz, const int32_t* rowptr,
                           const int32_t* colidx, const float* vals, float alpha, const float* B,
                           int32_t ldb, int32_t r, float* Y, int32_t ldy, int64_t chunk_nnz);
/* The same product with the rows visited in the order row_order[0..m) (a permutation of the rows;
 * Y is identical). No counterpart in the reference: with rows sorted by descending length
 * (mmsb_csr_row_order) the longest rows start first instead of forming the kernel's tail. */
int mmsb_spmm_f32_ordered(void* stream, int32_t m, int32_t n, int64_t nnz, const int32_t* rowptr,
                          const int32_t* colidx, const float* vals, float alpha, const float* B,
                          int32_t ldb, int32_t r, float* Y, int32_t ldy, const int32_t* row_order);
/* order[0..m) = the rows of the device CSR matrix rowptr[m+1] sorted by descending length, ties in
 * ascending row order. Allocates and frees its temporaries; returns when the order is complete. */
int mmsb_csr_row_order(void* stream, int32_t m, const int32_t* rowptr, int32_t* order);


/*
 * Coefficients (c3, c1, c0) of the empirical third moment
 *   T = c3 E[y_A x y_B x y_C] + c1 (E[y_A x y_B x mu_C] + E[y_A x mu_B x y_C] + E[mu_A x y_B x y_C])
 *       + c0 mu_A x mu_B x mu_C
 * scaling MMSB_M2_GPU: (1, -a0/(a0+2), 2a0^2/((a0+1)(a0+2))), the 4x4 matrices of init_buff
 * (stpm.c:1026-1038); MMSB_M2_PAPER: ((a0+1)(a0+2)/2, -a0(a0+1)/2, a0^2) (The_second_term,
 * CPU Util.cpp:289-298).
 */
int mmsb_tensor_coeffs(double alpha0, int32_t scaling, double* out3);

/*
 * Loss and exact adjoint of the implicit tensor objective over one minibatch of nb nodes:
 *   L = theta/2 sum_ij <phi_i,phi_j>^3 - sum_i T_b(phi_i, phi_i, phi_i)
 *   grad = 3 theta Phi (G o G) - [T(I,phi,phi) + T(phi,I,phi) + T(phi,phi,I)],   G = Phi^T Phi
 * Replaces the reference's analytic gradient Diff_Loss / The_second_term (CPU Util.cpp:257-304),
 * which keeps only 3 T(phi,phi,I) and so equals this adjoint for symmetric inputs only; the GPU
 * update of stpm.c:649-691 is that expression / 3. Device pointers: y_* [nb x ldy] row-major
 * (FP64 or FP32), mu_* [k], phi [k x k] row-major (column i = eigenvector i), coeffs[3] (host),
 * grad [k x k], loss [1]; work holds mmsb_stgd_grad_workspace(k, nb) bytes. Allocates nothing.
 * Accumulation in FP64.
 */
size_t mmsb_stgd_grad_workspace(int32_t k, int32_t nb);
int mmsb_stgd_grad_f64(void* stream, int32_t k, int32_t nb, const double* ya, const double* yb, const double* yc,
                       int32_t ldy, const double* mu_a, const double* mu_b, const double* mu_c, const double* phi,
                       const double* coeffs, double theta, double* grad, double* loss, void* work);
int mmsb_stgd_grad_f32(void* stream, int32_t k, int32_t nb, const float* ya, const float* yb, const float* yc,
                       int32_t ldy, const double* mu_a, const double* mu_b, const double* mu_c, const double* phi,
                       const double* coeffs, double theta, double* grad, double* loss, void* work);
/* mmsb_stgd_grad_f32 without the workspace argument: its scratch is a fixed 32 MiB device array of
 * the library (one per GPU, kept until the process ends), so the call allocates nothing -- for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). No counterpart in the reference. k, nb whose mmsb_stgd_grad_workspace
 * exceeds 32 MiB are rejected with MMSB_ERR_ARG (e.g. k = 100 allows nb up to 6504, k = 256 up to
 * 2216). Calls on one GPU share its array: calls that overlap on one GPU (other streams or threads)
 * race and must be serialised by the caller. */
int mmsb_stgd_grad_f32_noalloc(void* stream, int32_t k, int32_t nb, const float* ya, const float* yb, const float* yc,
                               int32_t ldy, const double* mu_a, const double* mu_b, const double* mu_c, const double* phi,
                               const double* coeffs, double theta, double* grad, double* loss);

/* ------------------------------------------------------------------------------------------
 * Host graph containers and the seeded MMSB generator
 * ------------------------------------------------------------------------------------------ */

/* CSR adjacency block in host memory, owned by the library (freed by mmsb_graph_free).
 * Replaces the dense NA x NX double buffers that read_G fills (pre_proc.c:29,49-80). */
typedef struct mmsb_csr {
  int32_t n_rows, n_cols;
  int64_t nnz;
  int32_t* rowptr;  /* [n_rows + 1] */
  int32_t* colidx;  /* [nnz], ascending within a row */
  float* vals;      /* [nnz] or NULL for a binary graph (all ones) */
} mmsb_csr;

/* The five partition blocks the reference reads (stpm.h:63-69): G_XA, G_XB, G_XC (rows = X
 * nodes) and G_BA (rows = B), G_CA (rows = C), all with columns in the partner partition. */
typedef struct mmsb_graph {
  mmsb_csr xa, xb, xc, ba, ca;
  int64_t n_nodes, n_edges;
  int32_t k;
  /* ground truth, filled only when requested (global node order): */
  double* theta;    /* [n_nodes] degree propensities                      */
  int8_t* part;     /* [n_nodes] 0..3 = X, A, B, C                          */
  int32_t* local;   /* [n_nodes] index inside the partition                 */
  double* pi;       /* [n_nodes * k] memberships, row-major                 */
} mmsb_graph;

enum { MMSB_GEN_AUTO = 0, MMSB_GEN_EXACT = 1, MMSB_GEN_BALLS = 2 };

typedef struct mmsb_gen_params {
  int32_t n_x, n_a, n_b, n_c, k;
  double alpha0;             /* Dirichlet concentration, pi ~ Dir(alpha0/k * 1)        */
  double p_in, p_out;        /* block matrix: p_in on the diagonal, p_out off it       */
  uint64_t seed;
  int32_t degree_corrected;  /* theta ~ Pareto(pareto_shape), normalised to mean 1     */
  double pareto_shape;
  int32_t mode;              /* MMSB_GEN_*: exact pairwise Bernoulli or ball dropping  */
  int32_t threads;           /* 0 = hardware concurrency (capped at 16)               */
  int32_t verbose;           /* progress lines on stdout, one per stage               */
} mmsb_gen_params;

/* ---- partition files (read_G, pre_proc.c:49-80; read_G_sparse, CPU IOfn.cpp:21-54) ---------- */

/* Repeated entries: MMSB_DUP_ONE = binary (entry = 1, BINARY pre_proc.c:67-68); MMSB_DUP_LAST =
 * last value wins (EXPECTED, pre_proc.c:64-65); MMSB_DUP_SUM = summed triplets (IOfn.cpp:51). */
enum { MMSB_DUP_ONE = 0, MMSB_DUP_LAST = 1, MMSB_DUP_SUM = 2 };

/* Reads a 1-based COO text file into a host CSR block (columns ascending). first_is_row selects
 * which column indexes rows; weighted reads a third value per entry; feof_repeat re-applies the
 * last entry as the reference's while(!feof) loops do. Status MMSB_ERR_IO if the file is missing
 * (reference: exit(1), pre_proc.c:54-58) or an index is out of range. */
int mmsb_read_coo(const char* path, int32_t n_rows, int32_t n_cols, int32_t first_is_row, int32_t weighted,
                  int32_t dup_policy, int32_t feof_repeat, mmsb_csr* out);
void mmsb_csr_free(mmsb_csr* m);
/* 1-based COO text writer (row<TAB>col[<TAB>value] per nonzero). */
int mmsb_write_coo(const char* path, const mmsb_csr* m);
/* One "%5.25e" value per line (write_mat, post_proc.c:45-59; stpm.c:926-951). */
int mmsb_write_values(const char* path, const double* v, int64_t n);
/* Nonzeros of a k x n matrix given node-major ([n][k]) as "row\tcol\tvalue" triplets in
 * column-major order, 6 significant digits (write_pi, CPU IOfn.cpp:57-67). */
int mmsb_write_pi_triplets(const char* path, const double* pi_nk, int64_t n, int32_t k);

/* Defaults: 1000 nodes per partition, k = 10, alpha0 = 1, p_in 0.6, p_out 0.02, seed 0. */
void mmsb_gen_params_default(mmsb_gen_params* p);
/* Seeded synthetic MMSB graph (no reference counterpart: the reference reads files only).
 * Bit-identical for a given parameter set whatever the thread count. */
int mmsb_generate(const mmsb_gen_params* p, int32_t keep_truth, mmsb_graph** out);
void mmsb_graph_free(mmsb_graph* g);


/* ------------------------------------------------------------------------------------------
 * Pipeline configuration (the stpm.h macros as a runtime struct) and the context handle
 * ------------------------------------------------------------------------------------------ */

enum { MMSB_PREC_FP32 = 0, MMSB_PREC_BF16 = 1 };
enum { MMSB_M2_PAPER = 0, MMSB_M2_GPU = 1 };
/* Neighbour vectors: MMSB_Y_STORED keeps Y_A, Y_B, Y_C (3 n_X x k, as pre_proc.c:339-345 does);
 * MMSB_Y_RECOMPUTE keeps only W, Z_B^T W, Z_C^T W and rebuilds each STGD minibatch's rows from
 * CSR before the step (checkpointed STGD; the whitening sketches and V are released after
 * whitening); MMSB_Y_AUTO recomputes when the stored form would not fit in 90% of the free
 * device memory. Both forms give identical STGD results. */
enum { MMSB_Y_AUTO = 0, MMSB_Y_STORED = 1, MMSB_Y_RECOMPUTE = 2 };
/* Orthonormal renormalisation of Phi after a step: MMSB_RENORM_NS = ns_iters Newton-Schulz
 * iterations towards the polar factor (GEMMs only), logging ||Phi_next - Phi||^2; MMSB_RENORM_QR =
 * the reference's QR (culaDeviceDgeqrfp + Dorgqr, stpm.c:774-786) as a k x k CholeskyQR2 (FP64 Gram and
 * Cholesky, R with a positive diagonal), logging ||Phi~ - Phi||^2 before it as stpm.c:705-743 does. */
enum { MMSB_RENORM_NS = 0, MMSB_RENORM_QR = 1 };

typedef struct stpm_config {
  int32_t nx, na, nb, nc;     /* NX, NA, NB, NC (stpm.h:72-75)                                   */
  int32_t ncom;               /* NCOM = k (stpm.h:76)                                            */
  double alpha_0;             /* ALPHA_0 (stpm.h:78)                                             */
  double learn_rate;          /* LEARN_RATE = eta (stpm.h:79)                                    */
  double thresh;              /* THRESH: the loop ends after the step t (from 0) whose ||dPhi||_F^2 < thresh ... (stpm.h:94) */
  double pinv_tol;            /* PINV_TOL (stpm.h:95), relative to the largest value here        */
  int64_t sgd_iter_min;       /* ... when t > SGD_ITER_MIN (stpm.h:96, stpm.c:747); <= 0: no convergence test */
  int64_t sgd_iter_max;       /* or after the first step t > SGD_ITER_MAX (stpm.h:97, stpm.c:761), i.e. at most
                                 SGD_ITER_MAX + 2 steps; <= 0: no cap. Both are checked after every minibatch
                                 step of mmsb_stgd(ctx, -1, ...), within the configured epochs              */
  int32_t binary;             /* 1 = BINARY two-column files, 0 = EXPECTED three-column (stpm.h:56-57) */
  int32_t synthetic;          /* SYNTHETIC: G_BA/G_CA present, Pi_A/B/C produced (stpm.h:55)    */
  char folder_name[512];      /* FOLDER_NAME (stpm.h:47), with trailing '/'                       */
  char file_prefix[64];       /* "DBLP_" of FILE_A..FILE_CA (stpm.h:63-69)                        */
  char alpha_0_str[32];       /* ALPHA_0 as spelled in output file names (stpm.h:82-91)          */
  char learn_rate_str[32];    /* LEARN_RATE as spelled in output file names                      */
  /* extensions of the reference configuration */
  int32_t minibatch;          /* nodes per STGD step (reference: 1); rounded up to 64             */
  int32_t epochs;             /* passes over X                                                   */
  double theta;               /* weight of the orthogonality term (reference GPU: 1)            */
  int32_t renorm_every;       /* orthonormal renormalisation every n steps (reference: 1), 0 = never */
  int32_t ns_iters;           /* Newton-Schulz iterations per renormalisation                    */
  int32_t ns_degree;          /* 3: X <- X (3I - G)/2; 5: X <- X (15I - 10G + 3G^2)/8 (G = X^T X)  */
  int32_t precision;          /* MMSB_PREC_*: STGD contraction precision                         */
  int32_t polish_epochs;      /* FP32 epochs that follow BF16 epochs                             */
  int32_t oversample;         /* r - k of the randomized range finders                           */
  int32_t power_iters;        /* power iterations of the range finders                           */
  int32_t m2_scaling;         /* MMSB_M2_PAPER (CPU Util.cpp:81,109-113) / MMSB_M2_GPU (pre_proc.c:292-302) */
  double member_tau;          /* membership threshold on the normalised Pi (reference: 0)        */
  uint64_t seed;              /* replaces time(NULL) / clock() seeding                           */
  int32_t ba_first_is_row;    /* Gb_a/Gc_a orientation: 1 = first column is the B/C node (CPU order) */
  int32_t feof_repeat;        /* re-apply the last entry of each input file (while(!feof) quirk) */
  int32_t device;             /* HIP device ordinal                                              */
  int32_t verbose;            /* progress lines on stdout                                        */
  int32_t use_graphs;         /* replay whole STGD epochs as captured HIP graphs                 */
  int32_t y_mode;             /* MMSB_Y_*: neighbour vectors stored or recomputed per minibatch  */
  int32_t renorm_mode;        /* MMSB_RENORM_*: Newton-Schulz (default) or the reference's QR    */
  int32_t csr_host;           /* 1: the CSR blocks (and their transposes) stay in host memory and are
                               * streamed to the device in row chunks for every product -- for graphs
                               * whose blocks do not fit in HBM; the neighbour vectors must then be
                               * stored (y_mode auto or stored). 0 (default): blocks resident in HBM.
                               * Results are identical either way.                                */
} stpm_config;

/* Fills the reference defaults of stpm.h:47-106 (DBLP sizes, NCOM 100, ALPHA_0 1, LEARN_RATE 1e-3,
 * THRESH 1, PINV_TOL 1e-6, BINARY, SYNTHETIC) plus: minibatch 1024, 1 epoch, theta 1, renorm every
 * step with 2 degree-3 Newton-Schulz iterations, FP32, oversampling 10, 1 power iteration, paper M2
 * scaling, tau 0, seed 0, CPU file orientation. */
void stpm_config_default(stpm_config* cfg);
/* sizeof(stpm_config) as compiled into the library, so that a binding that lays the struct out
 * itself (ctypes, cgo, JNI) can check its layout. No counterpart in the reference (stpm.h is a set of
 * macros there). */
size_t stpm_config_size(void);

typedef struct mmsb_ctx mmsb_ctx;

enum { MMSB_BLOCK_XA = 0, MMSB_BLOCK_XB = 1, MMSB_BLOCK_XC = 2, MMSB_BLOCK_BA = 3, MMSB_BLOCK_CA = 4 };

/* Quantities readable with mmsb_get (copied to host as double, row-major):                   */
enum {
  MMSB_Q_MU_A = 0, MMSB_Q_MU_B, MMSB_Q_MU_C,       /* [n_A], [n_B], [n_C] column means         */
  MMSB_Q_W,                                         /* [n_A x k] whitening matrix               */
  MMSB_Q_U, MMSB_Q_V, MMSB_Q_SIGMA,                 /* rank-k SVD of P_BC: [n_B x k], [n_C x k], [k] */
  MMSB_Q_LB, MMSB_Q_LC,                             /* Z_B = L_B U^T, Z_C = L_C V^T: [n_A x k]  */
  MMSB_Q_LAM_M2,                                    /* [k] top eigenvalues of M2                */
  MMSB_Q_YA, MMSB_Q_YB, MMSB_Q_YC,                  /* [n_X x k] whitened neighbour vectors     */
  MMSB_Q_MUW_A, MMSB_Q_MUW_B, MMSB_Q_MUW_C,         /* [k] whitened means                       */
  MMSB_Q_PHI,                                       /* [k x k] Phi, column i = eigenvector i    */
  MMSB_Q_FROB,                                      /* [steps] ||dPhi||_F^2 per step            */
  MMSB_Q_LAMBDA, MMSB_Q_LAMBDA_P, MMSB_Q_ALPHA_HAT, /* [k] eigenvalues, Lambda', alpha_hat      */
  MMSB_Q_PI_X, MMSB_Q_PI_A, MMSB_Q_PI_B, MMSB_Q_PI_C,     /* [n_p x k] unnormalised Pi          */
  MMSB_Q_PIN_X, MMSB_Q_PIN_A, MMSB_Q_PIN_B, MMSB_Q_PIN_C, /* [n_p x k] normalised Pi            */
  MMSB_Q_COUNT_X, MMSB_Q_COUNT_A, MMSB_Q_COUNT_B, MMSB_Q_COUNT_C, /* [n_p] memberships per node */
  MMSB_Q_TIMES,                                     /* [4] seconds: whiten, stgd, post, total   */
  MMSB_Q_SWEEPS,                                    /* [2] Jacobi sweeps of the last whitening's
                                                       r x r SVD and eigensolve                  */
  MMSB_Q__COUNT
};

/* Creates a context on cfg->device. Device memory is allocated lazily by each stage. */
int mmsb_create(const stpm_config* cfg, mmsb_ctx** out);
void mmsb_destroy(mmsb_ctx* ctx);
/* Uploads one host CSR block (copied; the caller keeps ownership). Transposes are built here. */
int mmsb_set_block(mmsb_ctx* ctx, int32_t which, const mmsb_csr* blk);
/* Reads FILE_A..FILE_CA of the configuration (read_G calls of pre_proc.c:146-148 and
 * post_proc.c:79-82; G_CA is read from FILE_CA, not FILE_BA as post_proc.c:82 does). */
int mmsb_read_inputs(mmsb_ctx* ctx);
/* Whitening and whitened neighbour vectors: the computation of pre_proc (pre_proc.c:151-366). */
int mmsb_whiten(mmsb_ctx* ctx);
/* Sets Phi0 (k x k host, column i = eigenvector i); otherwise the seeded orthonormal start. */
int mmsb_set_phi(mmsb_ctx* ctx, const double* phi);
/* Supplies y_A, y_B, y_C ([n_X x k] host) and whitened means ([k]) directly, bypassing whitening. */
int mmsb_set_yvectors(mmsb_ctx* ctx, int64_t n_x, int32_t k, const double* ya, const double* yb, const double* yc,
                      const double* mua, const double* mub, const double* muc);
/* STGD (the loop of GPU_Task, stpm.c:497-790): `steps` minibatch steps in `precision`; steps < 0
 * runs the configured epochs (BF16 epochs followed by polish_epochs in FP32 when precision is
 * BF16), ended after the first step that meets the THRESH / SGD_ITER_MIN / SGD_ITER_MAX test of
 * stpm.c:747-765 (evaluated on the device after every step, so captured epochs keep running
 * without host round trips). The tensor coefficients follow m2_scaling: the paper M2 with the
 * paper T (CPU Util.cpp:289-298), the GPU M2 with init_buff's T (stpm.c:1026-1038). */
int mmsb_stgd(mmsb_ctx* ctx, int64_t steps, int32_t precision);
/* Changes LEARN_RATE (stpm.h:79, a compile-time constant in the reference) for the following
 * mmsb_stgd calls of this context, e.g. a different rate for polishing epochs. */
int mmsb_set_learn_rate(mmsb_ctx* ctx, double learn_rate);
/* Asks a running (or the next) mmsb_stgd of this context to stop at the next epoch boundary, as
 * the reference's forceExit_flag (stpm.c:61, tested at stpm.c:754, set from the console at
 * stpm.c:1198) ends the STGD loop: that call returns MMSB_OK with Phi after its last completed
 * epoch (at least one), and eigenvalues / post_proc proceed as after convergence. Safe to call from
 * another thread; the request is consumed when mmsb_stgd returns. */
int mmsb_request_stop(mmsb_ctx* ctx);
/* Measurement entry (no counterpart in the reference): `steps` STGD steps as mmsb_stgd, launched
 * directly (never from a captured graph) with HIP events on the context stream around each phase;
 * phase_ms[0..2] receive the summed device time of the P1 kernel (Y Phi contractions and
 * (Phi^T Phi)^2), the P2 kernel (gradient step) and the renormalisation kernels, each interval
 * less the time between two events recorded back to back (measured in the same call). */
int mmsb_stgd_timed(mmsb_ctx* ctx, int64_t steps, int32_t precision, double* phase_ms);
/* Many independent instances (contexts with their neighbour vectors, on one device, each at an
 * epoch boundary, sharing ns_iters and renorm_every; n, k, minibatch, ALPHA_0 and LEARN_RATE may
 * differ) advanced together: every epoch of every member runs either in ONE persistent
 * (cooperative) launch, or -- after mmsb_batch_autotune finds it faster -- as one grouped launch per
 * phase slot replayed from a captured HIP graph. No counterpart in the reference, which runs one
 * instance per process.
 * The members' Phi, frob logs and step counts advance exactly as with mmsb_stgd(ctx, epochs * spe,
 * precision); a member driven by other calls in between must be put in a new batch. */
typedef struct mmsb_batch mmsb_batch;
int mmsb_batch_create(mmsb_ctx* const* ctxs, int32_t count, int32_t precision, mmsb_batch** out);
int mmsb_batch_run(mmsb_batch* batch, int32_t epochs);
/* Times one epoch for each candidate launch shape -- the persistent kernel at several grid sizes
 * (workgroups), and one grouped launch per phase slot replayed from a captured HIP graph -- restoring
 * every member's Phi afterwards, and keeps the fastest for later runs. best_grid receives the
 * persistent grid, or 0 when the grouped launches won. */
int mmsb_batch_autotune(mmsb_batch* batch, int32_t* best_grid, double* best_ms);
void mmsb_batch_destroy(mmsb_batch* batch);
/* Loss and adjoint (mmsb_stgd_grad_f32) over X rows [row0, row0 + nrows) for a host Phi [k x k].
 * recompute != 0 rebuilds those whitened rows from CSR (y = G_X* rows times W, Z_B^T W, Z_C^T W)
 * instead of reading the stored neighbour vectors -- the memory-saving form for large n k. */
int mmsb_grad(mmsb_ctx* ctx, int64_t row0, int32_t nrows, const double* phi, int32_t scaling, double theta,
              int32_t recompute, double* grad, double* loss);
/* Batch implicit tensor power method (CPU tensorDecom_alpha0, Util.cpp:131-206) on the neighbour
 * vectors (resident, or rebuilt from CSR chunk by chunk in the recompute form), as an alternative to STGD: `iters` iterations of phi_j <- T(I, phi_j, phi_j)
 * over all of X with the paper coefficients, lambda_j = ||.||, column normalisation, and, when
 * orthogonalize != 0, the symmetric orthonormalisation Phi (Phi^T Phi)^-1/2 the reference lacks
 * (without it the columns collapse onto the dominant component). Starts from the current Phi
 * (or the seeded Phi0); leaves Phi in the context and lambda in *lambda (host, k values). */
int mmsb_power_method(mmsb_ctx* ctx, int32_t iters, int32_t orthogonalize, double* lambda);
/* Eigenvalues, Lambda' and alpha_hat (stpm.c:796-870). */
int mmsb_eigen(mmsb_ctx* ctx);
/* Pi_X, Pi_A, Pi_B, Pi_C, normalisation and membership counts (post_proc.c:62-119). */
int mmsb_post(mmsb_ctx* ctx);
/* Writes Pi_*, evecs_, evals_, frob_norm_ and sim_info_ files named as stpm.h:82-91 (sim_info line for
 * line as stpm.c:957-975, svd1 / svd2 timing the randomized SVD of P_BC and the M2 eigenproblem), plus
 * the CPU program's pi_ and white_ triplet files of mmsb_write_triplets under the same naming. */
int mmsb_write_outputs(mmsb_ctx* ctx);
/* The CPU program's result files (CommunityDetection.cpp:121-144, write_pi IOfn.cpp:57-67): pi_path
 * receives the normalised memberships of all nodes, partitions concatenated X, A, B, C (k x N, each
 * node's normProbVector), white_path the n_A x k whitening matrix W; both as "row\tcol\tvalue"
 * triplets (1-based) of the nonzeros in column-major order, 6 significant digits. NULL skips a file. */
int mmsb_write_triplets(mmsb_ctx* ctx, const char* pi_path, const char* white_path);
/* Copies a quantity to host memory as double; count is the number of doubles available. */
int mmsb_get(mmsb_ctx* ctx, int32_t what, double* host, int64_t count);
/* Rows [row0, row0 + nrows) of a row-indexed quantity (tall matrices, vectors, counts). Y_A, Y_B
 * and Y_C of a context that recomputes its minibatches are rebuilt from CSR for the call. */
int mmsb_get_rows(mmsb_ctx* ctx, int32_t what, int64_t row0, int64_t nrows, double* host, int64_t count);
/* Copies an FP32-resident quantity (tall matrices: W, U, V, L_B, L_C, Y_A, Y_B, Pi_*) in its
 * native precision, compact row-major [rows x k]. */
int mmsb_get_f32(mmsb_ctx* ctx, int32_t what, float* host, int64_t count);
/* Number of doubles mmsb_get(what) writes (0 if not computed yet). */
int64_t mmsb_get_size(mmsb_ctx* ctx, int32_t what);
/* Blocks until all work of the context has finished. */
int mmsb_sync(mmsb_ctx* ctx);
/* Releases the process-wide resources the library keeps between calls -- the pinned staging of large
 * copies, the host-CSR SpMM's staging, slots and copy stream, the CSR sort temporaries -- and the
 * context of the program-level entry points (pre_proc / GPU_Task / post_proc). No counterpart in the
 * reference. Call it before the HIP runtime is torn down (the Python binding does so at interpreter
 * exit); mmsb_destroy of the last live context releases the same resources. Later calls re-create
 * them on demand. */
void mmsb_release(void);
/* Largest gathered matrix (bytes, n x ldb FP32) for which the SpMM kernels cut rows wider than 256
 * columns into 256-column strips (default 8 GiB; bytes < 0 restores it); returns the previous limit.
 * Process-wide; the results do not depend on it. No counterpart in the reference. */
int64_t mmsb_set_strip_limit(int64_t bytes);

/* ---- program-level drop-in (stpm.c / pre_proc.c / post_proc.c) -------------------------- */
/* Sets the configuration that pre_proc / GPU_Task / post_proc use (the stpm.h macros). */
int stpm_set_config(const stpm_config* cfg);
/* pre_proc.c:112: reads the three X blocks and whitens. Returns 0, or a negative status
 * (the reference exits with 1 on I/O and 2 on CULA errors). */
int pre_proc(void);
/* stpm.c:132: STGD, eigenvalues, post_proc and the output files; returns NULL on success or a
 * pointer to the (negative) status value on failure. */
void* GPU_Task(void* arg);
/* post_proc.c:62: Pi_X..Pi_C and their files (reference return type void, declared int at stpm.c:57). */
int post_proc(void);
/* forceExit_flag = T_TRUE (stpm.c:1198): the STGD loop of a running GPU_Task ends at the next epoch
 * boundary and the task goes on to the eigenvalues, post_proc and the output files. */
void stpm_force_exit(void);

#ifdef __cplusplus
}
#endif

#endif /* MMSB_STPM_H */
