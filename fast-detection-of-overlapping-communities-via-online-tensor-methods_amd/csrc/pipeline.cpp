// Host orchestration of the MMSB tensor-method pipeline on one GPU: whitening, whitened neighbour
// vectors, STGD, eigenvalues and Pi, behind the stpm.h-shaped configuration.
//
// The reference keeps every operand in process-global dense FP64 buffers sized at compile time
// (pre_proc.c:29-42, stpm.c:40-47, post_proc.c:34-42) and moves tiles over PCIe for each product
// (culaBlockDgemm, gemm_pinv_nys.c:27-145). Here a context owns device buffers: the adjacency
// blocks in CSR (and the CSR of their transposes), tall n x r FP32 sketches, small r x r FP64
// factors, and the STGD state; all products run in HBM on one stream.
//
// Whitening (pre_proc.c:151-366, with the intended mathematics of the Nystrom helpers):
//   mu_* = column means                                             (pre_proc.c:154-162)
//   P_BC = G_XB^T G_XC / n_X ~ U diag(sigma) V^T  (randomized SVD)  (pinv_nys_asym, gemm_pinv_nys.c:325-421)
//   Z_B = P_AC P_BC^+ = L_B U^T,  L_B = G_XA^T G_XC V diag(1/sigma) / n_X   (pre_proc.c:215-218)
//   Z_C = P_AB P_CB^+ = L_C V^T,  L_C = G_XA^T G_XB U diag(1/sigma) / n_X   (pre_proc.c:269)
//   M2  = s_main (L_C Sigma L_B^T + L_B Sigma L_C^T)/2 - s_shift mu_A mu_A^T  (pre_proc.c:292-302,
//         CPU Util.cpp:81-113), applied as an operator; top-k eigenpairs by a randomized range
//         finder and Rayleigh-Ritz; W = U_k Lambda_k^-1/2           (nystrom_whitening, gemm_pinv_nys.c:207-322)
//   y_A = W^T g_A, y_B = W^T Z_B g_B, y_C = W^T Z_C g_C             (pre_proc.c:339-366)
// Orthonormal bases come from CholeskyQR2 (FP64 Gram, Cholesky and triangular inverse on the GPU),
// small SVD / eigenproblems from the FP64 one-sided Jacobi kernel.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <new>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dense.h"
#include "host_util.h"
#include "mmsb_stpm.h"
#include "post.h"
#include "stgd.h"
#include "xfer.h"

namespace mmsb {

#define HCHK(expr)                                                                                     \
  do {                                                                                                 \
    hipError_t e_ = (expr);                                                                            \
    if (e_ != hipSuccess)                                                                              \
      return ::mmsb::set_error(MMSB_ERR_HIP, "%s:%d %s -> %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_)); \
  } while (0)
#define TRY(expr)                                                                                      \
  do {                                                                                                 \
    int s_ = (expr);                                                                                   \
    if (s_ != MMSB_OK) return s_;                                                                      \
  } while (0)

static constexpr double kCholShift = 1e-14;

// A CSR block: resident in device memory, or (csr_host) kept in host memory and streamed to the
// device in row chunks by every product.
struct DevCsr {
  int32_t n_rows = 0, n_cols = 0;
  int64_t nnz = 0;
  int32_t* rowptr = nullptr;
  int32_t* colidx = nullptr;
  float* vals = nullptr;
  int32_t* order = nullptr;  // rows by descending length: the SpMM's schedule
  bool host = false;         // the arrays below hold the block; the device pointers stay null
  std::vector<int32_t> h_rowptr, h_colidx;
  std::vector<float> h_vals;
};

static int64_t r64(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

}  // namespace mmsb

using namespace mmsb;

struct mmsb_ctx {
  stpm_config cfg{};
  hipStream_t s = nullptr;
  int32_t k = 0, r = 0, kp = 0, rld = 0, rmax = 0, b = 0, bl = 0;
  int64_t n[4] = {0, 0, 0, 0};  // X, A, B, C
  int64_t nmax = 0, n_pad = 0;
  DevCsr blk[5], blkT[5];
  bool have[5] = {false, false, false, false, false};
  std::vector<void*> allocs;
  // whitening scratch
  float* T[4] = {nullptr, nullptr, nullptr, nullptr};
  double *gram = nullptr, *R1 = nullptr, *R1i = nullptr, *R2 = nullptr, *R2i = nullptr, *Rq = nullptr;
  double* work = nullptr;
  size_t work_bytes = 0;
  unsigned char* jwork = nullptr;
  int32_t* dstat = nullptr;  // [8] chol status, jacobi sweeps
  double *Ut = nullptr, *Vt = nullptr, *sig = nullptr, *sinv = nullptr, *sgn = nullptr, *shift = nullptr;
  unsigned long long* best = nullptr;
  double *P1 = nullptr, *P2 = nullptr, *mrow = nullptr, *S = nullptr, *E = nullptr, *lamv = nullptr,
         *lamis = nullptr;
  double* mu[3] = {nullptr, nullptr, nullptr};
  float *W = nullptr, *U = nullptr, *V = nullptr, *LB = nullptr, *LC = nullptr, *MB = nullptr, *MC = nullptr;
  double *LBW = nullptr, *LCW = nullptr;
  bool whitened = false;
  // neighbour vectors
  float *YA = nullptr, *YB = nullptr, *YCt = nullptr;
  double* muw[3] = {nullptr, nullptr, nullptr};
  bool have_y = false;       // Y_A, Y_B, Y_C^T resident for all of X
  bool y_window = false;     // checkpointed STGD: each step recomputes its minibatch rows from CSR into yw
  // [b x kp], [b x kp], [kp x b] in FP32 and in BF16: one set per precision, so that the padding
  // columns (k..kp) and rows of each stay zero in its own format
  float* yw[3] = {nullptr, nullptr, nullptr};
  bf16_bits* yw16[3] = {nullptr, nullptr, nullptr};
  // STGD
  float* phi[2] = {nullptr, nullptr};
  int cur = 0;
  float *phit = nullptr, *gt = nullptr, *phia = nullptr, *h = nullptr, *cgt = nullptr, *cs_part = nullptr, *mab = nullptr;
  unsigned long long* ns_rowsum = nullptr;
  float *mpart = nullptr, *nsq = nullptr;
  double* fpart = nullptr;
  bf16_bits *phi16[2] = {nullptr, nullptr}, *phit16 = nullptr, *phia16 = nullptr, *phiT16 = nullptr, *phitT16 = nullptr,
            *phiaT16 = nullptr, *gt16 = nullptr, *h16 = nullptr, *nsq16 = nullptr;
  double* frob = nullptr;
  int64_t frob_cap = 0, steps_done = 0;
  bf16_bits *YA16 = nullptr, *YB16 = nullptr, *YCt16 = nullptr;
  double* frob_epoch = nullptr;            // per-epoch log written by captured epochs
  // captured epochs, [precision][cur][end test on]
  hipGraphExec_t epoch_graph[2][2][2] = {{{nullptr, nullptr}, {nullptr, nullptr}}, {{nullptr, nullptr}, {nullptr, nullptr}}};
  StgdStop* stopctl = nullptr;             // device state of the per-step end test (stpm.c:747-765)
  float *phiT = nullptr, *phitT = nullptr, *phiaT = nullptr;
  bool y16_valid = false;
  bool have_phi = false;
  // stop request (forceExit_flag, stpm.c:61,754): set from any thread, honoured between epochs
  std::atomic<int> stop_req{0};
  // eigen / post
  double *lam = nullptr, *lamp = nullptr, *ahat = nullptr, *phil = nullptr, *LBU = nullptr;
  float *Upost = nullptr, *ZBtU = nullptr;
  float* Pi[4] = {nullptr, nullptr, nullptr, nullptr};
  float* PiN[4] = {nullptr, nullptr, nullptr, nullptr};
  int32_t* cnt[4] = {nullptr, nullptr, nullptr, nullptr};
  bool eig_done = false, post_done = false;
  double times[4] = {0, 0, 0, 0};
  // seconds of the two factorisations of the last whitening: the randomized SVD of P_BC and the
  // M2 eigenproblem (the svd1 / svd2 timers of sim_info, stpm.c:884-893,970-971)
  double t_svd[2] = {0, 0};

  void log(const char* fmt, ...) const {
    if (!cfg.verbose) return;
    va_list ap;
    va_start(ap, fmt);
    std::vprintf(fmt, ap);
    va_end(ap);
    std::fflush(stdout);
  }
};

namespace mmsb {

template <typename T>
static int dalloc(mmsb_ctx* c, T*& p, size_t count) {
  if (p) return MMSB_OK;
  void* q = nullptr;
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess)
    return set_error(MMSB_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
  HCHK(hipMemsetAsync(q, 0, bytes, c->s));
  c->allocs.push_back(q);
  p = static_cast<T*>(q);
  return MMSB_OK;
}

template <typename T>
static void dfree(mmsb_ctx* c, T*& p) {
  if (!p) return;
  (void)hipStreamSynchronize(c->s);
  auto it = std::find(c->allocs.begin(), c->allocs.end(), (void*)p);
  if (it != c->allocs.end()) c->allocs.erase(it);
  (void)hipFree(p);
  p = nullptr;
}

// Captured epoch graphs hold the step parameters and buffer addresses in their kernel arguments.
static int drop_epoch_graphs(mmsb_ctx* c) {
  HCHK(hipStreamSynchronize(c->s));
  for (auto& plane : c->epoch_graph)
    for (auto& row : plane)
      for (hipGraphExec_t& g : row)
        if (g) {
          HCHK(hipGraphExecDestroy(g));
          g = nullptr;
        }
  return MMSB_OK;
}

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---- CSR upload (with the transpose built on the host by a counting sort) -------------------

static int upload_csr(mmsb_ctx* c, DevCsr& d, int32_t nr, int32_t nc, int64_t nnz, const int32_t* rp, const int32_t* ci,
                      const float* v) {
  d.n_rows = nr;
  d.n_cols = nc;
  d.nnz = nnz;
  if (c->cfg.csr_host) {  // kept in host memory (a copy owned by the context)
    try {
      d.h_rowptr.assign(rp, rp + (size_t)nr + 1);
      d.h_colidx.assign(ci, ci + (size_t)nnz);
      if (v) d.h_vals.assign(v, v + (size_t)nnz);
    } catch (const std::bad_alloc&) {
      return set_error(MMSB_ERR_NOMEM, "set_block: out of host memory for a %lld-nonzero block", (long long)nnz);
    }
    d.host = true;
    return MMSB_OK;
  }
  TRY(dalloc(c, d.rowptr, (size_t)nr + 1));
  TRY(dalloc(c, d.colidx, (size_t)std::max<int64_t>(nnz, 1)));
  TRY(h2d(c->s, d.rowptr, rp, sizeof(int32_t) * ((size_t)nr + 1)));
  TRY(h2d(c->s, d.colidx, ci, sizeof(int32_t) * (size_t)nnz));
  if (v) {
    TRY(dalloc(c, d.vals, (size_t)std::max<int64_t>(nnz, 1)));
    TRY(h2d(c->s, d.vals, v, sizeof(float) * (size_t)nnz));
  }
  TRY(dalloc(c, d.order, (size_t)nr));
  return csr_row_order(c->s, nr, d.rowptr, d.order);
}

// G^T of a host-resident block: a counting sort over the columns in which the rows are visited in
// order, so each row of G^T lists its entries by ascending row of G -- the order of the device's
// stable sort. Threads take contiguous row ranges; thread th's entries of column j are placed after
// those of threads th' < th, which keeps that order.
static int transpose_csr_host(const DevCsr& g, DevCsr& t) {
  t.n_rows = g.n_cols;
  t.n_cols = g.n_rows;
  t.nnz = g.nnz;
  t.host = true;
  const bool w = !g.h_vals.empty();
  const int64_t nr = g.n_rows, nc = g.n_cols;
  const int nt = g.nnz < ((int64_t)1 << 20) ? 1 : (int)std::min<int64_t>(std::min(16, std::max(1, default_threads())), nr);
  std::vector<std::vector<int32_t>> next((size_t)nt);
  try {
    t.h_rowptr.assign((size_t)nc + 1, 0);
    t.h_colidx.resize((size_t)g.nnz);
    if (w) t.h_vals.resize((size_t)g.nnz);
    for (auto& v : next) v.assign((size_t)nc, 0);
  } catch (const std::bad_alloc&) {
    return set_error(MMSB_ERR_NOMEM, "set_block: out of host memory for a transpose");
  }
  parallel_chunks(nr, nt, [&](int th, int64_t lo, int64_t hi) {  // per-thread column counts
    int32_t* cnt = next[(size_t)th].data();
    for (int64_t e = g.h_rowptr[(size_t)lo]; e < g.h_rowptr[(size_t)hi]; ++e) ++cnt[g.h_colidx[(size_t)e]];
  });
  parallel_chunks(nc, nt, [&](int, int64_t lo, int64_t hi) {  // column totals
    for (int64_t j = lo; j < hi; ++j) {
      int32_t s = 0;
      for (int th = 0; th < nt; ++th) s += next[(size_t)th][(size_t)j];
      t.h_rowptr[(size_t)j + 1] = s;
    }
  });
  for (int64_t j = 0; j < nc; ++j) t.h_rowptr[(size_t)j + 1] += t.h_rowptr[(size_t)j];
  parallel_chunks(nc, nt, [&](int, int64_t lo, int64_t hi) {  // counts -> each thread's first slot
    for (int64_t j = lo; j < hi; ++j) {
      int32_t pos = t.h_rowptr[(size_t)j];
      for (int th = 0; th < nt; ++th) {
        const int32_t c = next[(size_t)th][(size_t)j];
        next[(size_t)th][(size_t)j] = pos;
        pos += c;
      }
    }
  });
  parallel_chunks(nr, nt, [&](int th, int64_t lo, int64_t hi) {  // scatter, rows in order
    int32_t* nx = next[(size_t)th].data();
    for (int64_t i = lo; i < hi; ++i)
      for (int32_t e = g.h_rowptr[(size_t)i]; e < g.h_rowptr[(size_t)i + 1]; ++e) {
        const int32_t pos = nx[g.h_colidx[(size_t)e]]++;
        t.h_colidx[(size_t)pos] = (int32_t)i;
        if (w) t.h_vals[(size_t)pos] = g.h_vals[(size_t)e];
      }
  });
  return MMSB_OK;
}

// G^T built on the device from the uploaded G.
static int transpose_csr(mmsb_ctx* c, const DevCsr& g, DevCsr& t) {
  if (g.host) return transpose_csr_host(g, t);
  t.n_rows = g.n_cols;
  t.n_cols = g.n_rows;
  t.nnz = g.nnz;
  TRY(dalloc(c, t.rowptr, (size_t)t.n_rows + 1));
  TRY(dalloc(c, t.colidx, (size_t)std::max<int64_t>(t.nnz, 1)));
  if (g.vals) TRY(dalloc(c, t.vals, (size_t)std::max<int64_t>(t.nnz, 1)));
  TRY(csr_transpose(c->s, g.n_rows, g.n_cols, g.nnz, g.rowptr, g.colidx, g.vals, t.rowptr, t.colidx, t.vals));
  TRY(dalloc(c, t.order, (size_t)t.n_rows));
  return csr_row_order(c->s, t.n_rows, t.rowptr, t.order);
}

// Nonzeros per streamed chunk of a host-resident block (16 MiB of column indices). The blocks were
// checked by mmsb_set_block, so the streamed products skip the scan.
static constexpr int64_t kHostChunkNnz = (int64_t)1 << 22;

static int spmm(mmsb_ctx* c, const DevCsr& g, float alpha, const float* B, int32_t ldb, int32_t r, float* Y, int32_t ldy) {
  if (g.host)
    return spmm_f32_host_csr(c->s, g.n_rows, g.n_cols, g.nnz, g.h_rowptr.data(), g.h_colidx.data(),
                             g.h_vals.empty() ? nullptr : g.h_vals.data(), alpha, B, ldb, r, Y, ldy, kHostChunkNnz, false);
  return spmm_f32(c->s, g.n_rows, g.n_cols, g.nnz, g.rowptr, g.colidx, g.vals, alpha, B, ldb, r, Y, ldy, g.order);
}

// Rows [row0, row0 + nrows) of a block times B into Y (Y's first row is row0's).
static int spmm_rows(mmsb_ctx* c, hipStream_t s, const DevCsr& g, int64_t row0, int32_t nrows, const float* B,
                     int32_t ldb, int32_t r, float* Y, int32_t ldy) {
  if (!g.host) return spmm_f32(s, nrows, g.n_cols, g.nnz, g.rowptr + row0, g.colidx, g.vals, 1.f, B, ldb, r, Y, ldy);
  const int32_t z0 = g.h_rowptr[(size_t)row0];
  std::vector<int32_t> rp((size_t)nrows + 1);
  for (int32_t j = 0; j <= nrows; ++j) rp[(size_t)j] = g.h_rowptr[(size_t)row0 + j] - z0;
  return spmm_f32_host_csr(s, nrows, g.n_cols, rp[(size_t)nrows], rp.data(), g.h_colidx.data() + z0,
                           g.h_vals.empty() ? nullptr : g.h_vals.data() + z0, 1.f, B, ldb, r, Y, ldy, kHostChunkNnz, false);
}

// (1 / n_X) x the row sums of a block's transpose (the column means), into the device vector out: on the
// device for a resident block, on the host -- in the same order and FP64 -- for a host-resident one.
static int col_means(mmsb_ctx* c, const DevCsr& t, double scale, double* out) {
  if (!t.host) return csr_rowsum(c->s, t.n_rows, t.rowptr, t.vals, scale, out);
  std::vector<double> m((size_t)t.n_rows);
  for (int32_t i = 0; i < t.n_rows; ++i) {
    double acc;
    if (!t.h_vals.empty()) {
      acc = 0.0;
      for (int32_t e = t.h_rowptr[(size_t)i]; e < t.h_rowptr[(size_t)i + 1]; ++e) acc += (double)t.h_vals[(size_t)e];
    } else {
      acc = (double)(t.h_rowptr[(size_t)i + 1] - t.h_rowptr[(size_t)i]);
    }
    m[(size_t)i] = scale * acc;
  }
  return h2d(c->s, out, m.data(), sizeof(double) * m.size());
}

// ---- CholeskyQR2: Q = X R^{-1} (Q may alias X); optionally R = R2 R1 -------------------------
// Y = X C for a well-conditioned r x r factor C (the second CholeskyQR pass, orthogonal factors):
// FP32 MFMA on tall matrices (the FP32 sums stay at the FP32 storage level of Y), the FP64-accumulating
// product otherwise.
static int nn_wellcond(mmsb_ctx* c, int64_t n, int32_t p, int32_t q, const float* X, int32_t ldx, const double* C,
                       int32_t ldc, float* Y, int32_t ldy, bool c_trans) {
  if (n >= 4096) return gemm_nn_f32(c->s, n, p, q, 1.0, X, ldx, C, ldc, 0.0, Y, ldy, c_trans);
  return gemm_nn<float, float>(c->s, n, p, q, 1.0, X, ldx, C, ldc, 0.0, Y, ldy, c_trans);
}

// A basis that only feeds the next product of a subspace iteration needs to span the right space
// with a well-conditioned basis, not to be orthonormal to working precision: one_pass stops after the
// first CholeskyQR pass (Q = X R1^-1, orthogonality error ~kappa(X)^2 eps64); the bases that are used
// as such (U, V, the M2 Rayleigh-Ritz basis, R for the SVD) take both passes.
static int orth(mmsb_ctx* c, float* X, int64_t n, int32_t cols, int32_t ld, float* Q, float* tmp, double* R_out,
                int status_slot, bool one_pass = false) {
  hipStream_t s = c->s;
  TRY((gemm_tn<float, float>(s, n, cols, cols, 1.0, X, ld, X, ld, 0.0, c->gram, c->work)));
  TRY(chol_inv(s, cols, c->gram, kCholShift, c->R1, c->R1i, c->dstat + status_slot));
  TRY((gemm_nn<float, float>(s, n, cols, cols, 1.0, X, ld, c->R1i, cols, 0.0, tmp, ld)));
  if (one_pass && !R_out) {  // the tiled product cannot run in place (Q may alias X): copy back
    HCHK(hipMemcpyAsync(Q, tmp, sizeof(float) * (size_t)n * ld, hipMemcpyDeviceToDevice, s));
    return MMSB_OK;
  }
  TRY((gemm_tn<float, float>(s, n, cols, cols, 1.0, tmp, ld, tmp, ld, 0.0, c->gram, c->work)));
  TRY(chol_inv(s, cols, c->gram, kCholShift, c->R2, c->R2i, c->dstat + status_slot + 1));
  TRY(nn_wellcond(c, n, cols, cols, tmp, ld, c->R2i, cols, Q, ld, false));  // R2 = I + O(kappa eps)
  if (R_out) TRY((gemm_nn<double, double>(s, cols, cols, cols, 1.0, c->R2, cols, c->R1, cols, 0.0, R_out, cols)));
  return MMSB_OK;
}

static int check_status(mmsb_ctx* c, const char* where) {
  int32_t st[8];
  HCHK(hipMemcpyAsync(st, c->dstat, sizeof(st), hipMemcpyDeviceToHost, c->s));
  HCHK(hipStreamSynchronize(c->s));
  for (int i = 0; i < 6; ++i)
    if (st[i] != 0) return set_error(MMSB_ERR_NUMERIC, "%s: Cholesky pivot %d not positive (slot %d)", where, st[i], i);
  return MMSB_OK;
}

// M2 applied to X [n_A x r]:  OUT = s_main/2 (L_C Sigma L_B^T X + L_B Sigma L_C^T X) - s_shift mu (mu^T X)
//                                  [+ s_shift diag(mu^2) X in the GPU scaling]
static int m2_apply(mmsb_ctx* c, const float* X, float* OUT) {
  hipStream_t s = c->s;
  const int64_t na = c->n[1];
  const int32_t k = c->k, r = c->r, kp = c->kp, ld = c->rld;
  const bool paper = c->cfg.m2_scaling == MMSB_M2_PAPER;
  const double a0 = c->cfg.alpha_0;
  const double s_main = paper ? a0 + 1.0 : 1.0;
  const double s_shift = paper ? a0 : a0 / (a0 + 1.0);
  // The range finder needs the operator's action to ~1e-6, not to FP64: on tall inputs its two
  // products L^T X and the two L P products run on FP32 MFMA (the r x r Rayleigh-Ritz matrix that
  // fixes the eigenvalues is formed in FP64 by the caller).
  const bool f32 = na >= 4096;
  if (f32) {
    TRY(gemm_tn_f32(s, na, k, r, 1.0, c->LB, kp, X, ld, 0.0, c->P1, c->work));
    TRY(gemm_tn_f32(s, na, k, r, 1.0, c->LC, kp, X, ld, 0.0, c->P2, c->work));
  } else {
    TRY((gemm_tn<float, float>(s, na, k, r, 1.0, c->LB, kp, X, ld, 0.0, c->P1, c->work)));
    TRY((gemm_tn<float, float>(s, na, k, r, 1.0, c->LC, kp, X, ld, 0.0, c->P2, c->work)));
  }
  TRY((gemm_tn<double, float>(s, na, 1, r, 1.0, c->mu[0], 1, X, ld, 0.0, c->mrow, c->work)));
  TRY(scale_rows(s, k, r, c->P1, r, c->sig, 0.5 * s_main));
  TRY(scale_rows(s, k, r, c->P2, r, c->sig, 0.5 * s_main));
  if (f32) {
    TRY(gemm_nn_f32(s, na, k, r, 1.0, c->LC, kp, c->P1, r, 0.0, OUT, ld));
    TRY(gemm_nn_f32(s, na, k, r, 1.0, c->LB, kp, c->P2, r, 1.0, OUT, ld));
  } else {
    TRY((gemm_nn<float, float>(s, na, k, r, 1.0, c->LC, kp, c->P1, r, 0.0, OUT, ld)));
    TRY((gemm_nn<float, float>(s, na, k, r, 1.0, c->LB, kp, c->P2, r, 1.0, OUT, ld)));
  }
  TRY((gemm_nn<double, float>(s, na, 1, r, -s_shift, c->mu[0], 1, c->mrow, r, 1.0, OUT, ld)));
  if (!paper) TRY(diag_axpy(s, na, r, c->mu[0], s_shift, X, ld, OUT, ld));
  return MMSB_OK;
}

static int alloc_small(mmsb_ctx* c) {
  const int32_t rm = c->rmax;
  TRY(dalloc(c, c->gram, (size_t)rm * rm));
  TRY(dalloc(c, c->R1, (size_t)rm * rm));
  TRY(dalloc(c, c->R1i, (size_t)rm * rm));
  TRY(dalloc(c, c->R2, (size_t)rm * rm));
  TRY(dalloc(c, c->R2i, (size_t)rm * rm));
  TRY(dalloc(c, c->Rq, (size_t)rm * rm));
  size_t wb = 0;
  for (int64_t nn : {c->nmax, (int64_t)rm})
    for (int32_t p : {rm, c->k, 1})
      for (int32_t q : {rm, c->k, 1}) wb = std::max(wb, tn_workspace_bytes(nn, p, q));
  TRY(dalloc(c, c->work, wb / sizeof(double) + 1));
  c->work_bytes = wb;
  TRY(dalloc(c, c->jwork, jacobi_workspace_bytes(rm)));
  TRY(dalloc(c, c->dstat, 8));
  TRY(dalloc(c, c->Ut, (size_t)rm * rm));
  TRY(dalloc(c, c->Vt, (size_t)rm * rm));
  TRY(dalloc(c, c->sig, rm));
  TRY(dalloc(c, c->sinv, rm));
  TRY(dalloc(c, c->sgn, rm));
  TRY(dalloc(c, c->shift, 1));
  TRY(dalloc(c, c->best, rm));
  TRY(dalloc(c, c->P1, (size_t)c->k * rm));
  TRY(dalloc(c, c->P2, (size_t)c->k * rm));
  TRY(dalloc(c, c->mrow, rm));
  TRY(dalloc(c, c->S, (size_t)rm * rm));
  TRY(dalloc(c, c->E, (size_t)rm * rm));
  TRY(dalloc(c, c->lamv, rm));
  TRY(dalloc(c, c->lamis, rm));
  return MMSB_OK;
}

static int alloc_whitening(mmsb_ctx* c) {
  TRY(alloc_small(c));
  for (int i = 0; i < 4; ++i) TRY(dalloc(c, c->T[i], (size_t)c->nmax * c->rld));
  for (int i = 0; i < 3; ++i) TRY(dalloc(c, c->mu[i], (size_t)c->n[i + 1]));
  TRY(dalloc(c, c->W, (size_t)c->n[1] * c->kp));
  TRY(dalloc(c, c->U, (size_t)c->n[2] * c->kp));
  TRY(dalloc(c, c->V, (size_t)c->n[3] * c->kp));
  TRY(dalloc(c, c->LB, (size_t)c->n[1] * c->kp));
  TRY(dalloc(c, c->LC, (size_t)c->n[1] * c->kp));
  TRY(dalloc(c, c->MB, (size_t)c->n[2] * c->kp));
  TRY(dalloc(c, c->MC, (size_t)c->n[3] * c->kp));
  TRY(dalloc(c, c->LBW, (size_t)c->k * c->k));
  TRY(dalloc(c, c->LCW, (size_t)c->k * c->k));
  return MMSB_OK;
}

// Bytes the resident form would add from whitening on: sketches and factors, Y_A/Y_B/Y_C^T with their
// BF16 copies, and the post_proc buffers.
static double resident_bytes(const mmsb_ctx* c) {
  const double kp4 = 4.0 * c->kp;
  const double whiten = 4.0 * (double)c->nmax * c->rld * 4.0 + (3.0 * c->n[1] + 2.0 * c->n[2] + 2.0 * c->n[3]) * kp4;
  const double y = 3.0 * (double)c->n_pad * c->kp * 6.0;
  const double post = (double)(c->n[1] + c->n[2]) * kp4 + 2.0 * (double)(c->n[0] + c->n[1] + c->n[2] + c->n[3]) * kp4;
  return whiten + y + post;
}

// MMSB_Y_AUTO: recompute the minibatches when the resident form would not fit in 90% of the free
// device memory.
static int choose_y_window(mmsb_ctx* c, bool& window) {
  const int32_t mode = c->cfg.y_mode;
  if (mode != MMSB_Y_AUTO) {
    window = mode == MMSB_Y_RECOMPUTE;
    if (window && c->cfg.csr_host)
      return set_error(MMSB_ERR_ARG, "whiten: recomputed minibatches read the CSR blocks on the device; with csr_host "
                                     "the neighbour vectors must be stored (y_mode auto or stored)");
    return MMSB_OK;
  }
  size_t free_b = 0, total_b = 0;
  HCHK(hipMemGetInfo(&free_b, &total_b));
  const double need = resident_bytes(c);
  window = need > 0.9 * (double)free_b;
  if (window && c->cfg.csr_host)
    return set_error(MMSB_ERR_NOMEM, "whiten: the stored neighbour vectors need %.1f GB of %.1f GB free, and with "
                                     "csr_host the minibatches cannot be recomputed from CSR", need * 1e-9,
                     (double)free_b * 1e-9);
  c->log("whiten: resident form needs %.1f GB of %.1f GB free: %s\n", need * 1e-9, (double)free_b * 1e-9,
         window ? "minibatches recomputed from CSR" : "neighbour vectors kept");
  return MMSB_OK;
}

static int alloc_window(mmsb_ctx* c) {
  for (float*& p : c->yw) TRY(dalloc(c, p, (size_t)c->b * c->kp));
  for (bf16_bits*& p : c->yw16) TRY(dalloc(c, p, (size_t)c->b * c->kp));
  return MMSB_OK;
}

// Enqueues the whitened rows [row0, row0 + nb) of X into the window buffers (FP32 or BF16).
static int fill_window(mmsb_ctx* c, bool bf16, int64_t row0, int32_t nb) {
  YRowsJobs jobs{};
  const float* R[3] = {c->W, c->MB, c->MC};
  for (int i = 0; i < 3; ++i) {
    const DevCsr& g = c->blk[i];
    void* out = bf16 ? (void*)c->yw16[i] : (void*)c->yw[i];
    jobs.j[i] = YRowsJob{g.rowptr + row0, g.colidx, g.vals, R[i], out, i == 2 ? (int64_t)c->b : (int64_t)c->kp, i == 2,
                         (int64_t)g.n_cols};
  }
  jobs.count = 3;
  return spmm_yrows(c->s, jobs, nb, c->k, c->kp, bf16);
}

static int alloc_y(mmsb_ctx* c) {
  TRY(dalloc(c, c->YA, (size_t)c->n_pad * c->kp));
  TRY(dalloc(c, c->YB, (size_t)c->n_pad * c->kp));
  TRY(dalloc(c, c->YCt, (size_t)c->kp * c->n_pad));
  for (int i = 0; i < 3; ++i) TRY(dalloc(c, c->muw[i], (size_t)c->kp));
  return MMSB_OK;
}

static int whiten_impl(mmsb_ctx* c) {
  hipStream_t s = c->s;
  const stpm_config& cf = c->cfg;
  const int32_t k = c->k, r = c->r, kp = c->kp, ld = c->rld;
  const int64_t nx = c->n[0], na = c->n[1], nb = c->n[2], ncn = c->n[3];
  const float inv = (float)(1.0 / (double)nx);
  const DevCsr &XA = c->blk[0], &XB = c->blk[1], &XC = c->blk[2];
  const DevCsr &XAt = c->blkT[0], &XBt = c->blkT[1], &XCt = c->blkT[2];
  float **T = c->T;
  bool window = false;
  TRY(choose_y_window(c, window));
  if (window != c->y_window) TRY(drop_epoch_graphs(c));
  c->y_window = window;
  TRY(alloc_whitening(c));
  if (window) {
    for (int i = 0; i < 3; ++i) TRY(dalloc(c, c->muw[i], (size_t)c->kp));
    TRY(alloc_window(c));
  } else {
    TRY(alloc_y(c));
  }
  // means (pre_proc.c:154-162)
  TRY(col_means(c, XAt, 1.0 / (double)nx, c->mu[0]));
  TRY(col_means(c, XBt, 1.0 / (double)nx, c->mu[1]));
  TRY(col_means(c, XCt, 1.0 / (double)nx, c->mu[2]));
  c->log("whiten: randomized SVD of P_BC (n_B=%lld n_C=%lld r=%d q=%d)\n", (long long)nb, (long long)ncn, r, cf.power_iters);
  HCHK(hipStreamSynchronize(s));
  double t_mark = now_s();
  // range of P_BC
  TRY(gaussian_fill(s, cf.seed, 0, ncn, r, T[0], ld));
  TRY(spmm(c, XC, 1.f, T[0], ld, r, T[1], ld));
  TRY(spmm(c, XBt, inv, T[1], ld, r, T[2], ld));
  TRY(orth(c, T[2], nb, r, ld, T[2], T[3], nullptr, 0, cf.power_iters > 0));
  for (int q = 0; q < cf.power_iters; ++q) {
    TRY(spmm(c, XB, 1.f, T[2], ld, r, T[1], ld));
    TRY(spmm(c, XCt, inv, T[1], ld, r, T[0], ld));
    TRY(orth(c, T[0], ncn, r, ld, T[0], T[3], nullptr, 2, true));
    TRY(spmm(c, XC, 1.f, T[0], ld, r, T[1], ld));
    TRY(spmm(c, XBt, inv, T[1], ld, r, T[2], ld));
    TRY(orth(c, T[2], nb, r, ld, T[2], T[3], nullptr, 0, q + 1 < cf.power_iters));
  }
  // P_CB Q = Q2 R2  =>  P_BC ~ Q R2^T Q2^T
  TRY(spmm(c, XB, 1.f, T[2], ld, r, T[1], ld));
  TRY(spmm(c, XCt, inv, T[1], ld, r, T[0], ld));
  TRY(orth(c, T[0], ncn, r, ld, T[0], T[3], c->Rq, 2));
  // SVD of R2^T: the row-major R2 buffer is R2^T in column-major order
  TRY(jacobi_svd(s, r, c->Rq, c->Ut, c->sig, c->Vt, c->jwork, c->dstat + 6));
  TRY(nn_wellcond(c, nb, r, k, T[2], ld, c->Ut, r, c->U, kp, true));   // U = Q Ut[:, :k] (Ut orthogonal)
  TRY(nn_wellcond(c, ncn, r, k, T[0], ld, c->Vt, r, c->V, kp, true));  // V = Q2 Vt[:, :k]
  TRY(col_sign<float>(s, nb, k, c->U, kp, c->best, c->sgn));
  TRY(scale_cols<float>(s, nb, k, c->U, kp, c->sgn));
  TRY(scale_cols<float>(s, ncn, k, c->V, kp, c->sgn));
  TRY(pinv_scale(s, k, c->sig, 0.0, cf.pinv_tol, 0, c->sinv, nullptr));
  HCHK(hipStreamSynchronize(s));
  c->t_svd[0] = now_s() - t_mark;
  // L_B, L_C
  c->log("whiten: Z_B / Z_C factors\n");
  TRY(spmm(c, XC, 1.f, c->V, kp, k, T[1], ld));
  TRY(spmm(c, XAt, inv, T[1], ld, k, c->LB, kp));
  TRY(scale_cols<float>(s, na, k, c->LB, kp, c->sinv));
  TRY(spmm(c, XB, 1.f, c->U, kp, k, T[1], ld));
  TRY(spmm(c, XAt, inv, T[1], ld, k, c->LC, kp));
  TRY(scale_cols<float>(s, na, k, c->LC, kp, c->sinv));
  // M2: range finder, Rayleigh-Ritz, W
  c->log("whiten: M2 eigenpairs (n_A=%lld)\n", (long long)na);
  HCHK(hipStreamSynchronize(s));
  t_mark = now_s();
  TRY(gaussian_fill(s, cf.seed, 1, na, r, T[0], ld));
  TRY(m2_apply(c, T[0], T[2]));
  TRY(orth(c, T[2], na, r, ld, T[2], T[3], nullptr, 4, cf.power_iters > 0));
  int qbuf = 2;
  for (int q = 0; q < cf.power_iters; ++q) {
    const int other = qbuf == 2 ? 0 : 2;
    TRY(m2_apply(c, T[qbuf], T[other]));
    TRY(orth(c, T[other], na, r, ld, T[other], T[3], nullptr, 4, q + 1 < cf.power_iters));
    qbuf = other;
  }
  TRY(m2_apply(c, T[qbuf], T[1]));
  TRY((gemm_tn<float, float>(s, na, r, r, 1.0, T[qbuf], ld, T[1], ld, 0.0, c->S, c->work)));
  TRY(sym_shift(s, r, c->S, c->shift));
  TRY(jacobi_svd(s, r, c->S, c->E, c->lamv, nullptr, c->jwork, c->dstat + 7));  // S SPD: eigenvectors = U
  double hshift = 0.0;
  HCHK(hipMemcpyAsync(&hshift, c->shift, sizeof(double), hipMemcpyDeviceToHost, s));
  HCHK(hipStreamSynchronize(s));
  TRY(pinv_scale(s, r, c->lamv, hshift, cf.pinv_tol, 1, c->lamis, c->lamv));  // lamv <- sigma - shift
  TRY(nn_wellcond(c, na, r, k, T[qbuf], ld, c->E, r, c->W, kp, true));  // E orthogonal
  TRY(scale_cols<float>(s, na, k, c->W, kp, c->lamis));
  TRY(col_sign<float>(s, na, k, c->W, kp, c->best, c->sgn));
  TRY(scale_cols<float>(s, na, k, c->W, kp, c->sgn));
  HCHK(hipStreamSynchronize(s));
  c->t_svd[1] = now_s() - t_mark;
  // neighbour vectors (pre_proc.c:339-366)
  // y rows are G_X* rows times W, Z_B^T W = U (L_B^T W) and Z_C^T W = V (L_C^T W)
  c->log(window ? "whiten: neighbour-vector factors (rows recomputed per STGD minibatch)\n" : "whiten: neighbour vectors\n");
  TRY((gemm_tn<float, float>(s, na, k, k, 1.0, c->LB, kp, c->W, kp, 0.0, c->LBW, c->work)));
  TRY((gemm_nn<float, float>(s, nb, k, k, 1.0, c->U, kp, c->LBW, k, 0.0, c->MB, kp)));
  TRY((gemm_tn<float, float>(s, na, k, k, 1.0, c->LC, kp, c->W, kp, 0.0, c->LCW, c->work)));
  TRY((gemm_nn<float, float>(s, ncn, k, k, 1.0, c->V, kp, c->LCW, k, 0.0, c->MC, kp)));
  if (!window) {
    TRY(spmm(c, XA, 1.f, c->W, kp, k, c->YA, kp));
    TRY(spmm(c, XB, 1.f, c->MB, kp, k, c->YB, kp));
    TRY(spmm(c, XC, 1.f, c->MC, kp, k, T[1], ld));
    TRY(transpose_f32(s, nx, k, T[1], ld, c->YCt, c->n_pad));
  }
  TRY((gemm_tn<float, double>(s, na, k, 1, 1.0, c->W, kp, c->mu[0], 1, 0.0, c->muw[0], c->work)));
  TRY((gemm_tn<float, double>(s, nb, k, 1, 1.0, c->MB, kp, c->mu[1], 1, 0.0, c->muw[1], c->work)));
  TRY((gemm_tn<float, double>(s, ncn, k, 1, 1.0, c->MC, kp, c->mu[2], 1, 0.0, c->muw[2], c->work)));
  TRY(check_status(c, "whitening"));
  c->whitened = true;
  c->have_y = !window;
  c->y16_valid = false;
  if (window) {  // what STGD and post_proc do not read goes back to the device: the sketches and V
    for (float*& p : c->T) dfree(c, p);
    dfree(c, c->V);
  }
  return MMSB_OK;
}

// ---- STGD --------------------------------------------------------------------------------------
static int alloc_stgd(mmsb_ctx* c) {
  const size_t kk = (size_t)c->kp * c->kp;
  TRY(dalloc(c, c->phi[0], kk));
  TRY(dalloc(c, c->phi[1], kk));
  TRY(dalloc(c, c->phit, kk));
  TRY(dalloc(c, c->gt, kk));
  TRY(dalloc(c, c->phia, kk));
  TRY(dalloc(c, c->h, kk));
  TRY(dalloc(c, c->cgt, (size_t)c->kp * c->b));
  TRY(dalloc(c, c->cs_part, (size_t)(c->b / 32) * c->kp));
  TRY(dalloc(c, c->mab, (size_t)2 * c->kp));
  TRY(dalloc(c, c->ns_rowsum, (size_t)2 * c->kp));
  TRY(dalloc(c, c->mpart, (size_t)(c->kp / 32) * 2 * c->kp));
  TRY(dalloc(c, c->fpart, (size_t)(c->kp / 32) * (c->kp / 32)));
  TRY(dalloc(c, c->phiT, kk));
  TRY(dalloc(c, c->phitT, kk));
  TRY(dalloc(c, c->phiaT, kk));
  TRY(dalloc(c, c->nsq, kk));
  return MMSB_OK;
}

static int init_phi(mmsb_ctx* c) {
  // Phi0 = CholeskyQR2 of a seeded k x k Gaussian (replaces the time-seeded QR of rand(),
  // stpm.c:1041-1064); phit and phia serve as scratch.
  TRY(alloc_small(c));
  TRY(gaussian_fill(c->s, c->cfg.seed, 2, c->k, c->k, c->phit, c->kp));
  TRY(orth(c, c->phit, c->k, c->k, c->kp, c->phi[c->cur], c->phia, nullptr, 0));
  TRY(check_status(c, "Phi0"));
  c->have_phi = true;
  return MMSB_OK;
}

static int prepare_bf16(mmsb_ctx* c) {
  const size_t ny = (size_t)c->n_pad * c->kp;
  if (c->have_y) {
    TRY(dalloc(c, c->YA16, ny));
    TRY(dalloc(c, c->YB16, ny));
    TRY(dalloc(c, c->YCt16, ny));
  }
  const size_t kk2 = (size_t)2 * c->kp * c->kp;
  for (bf16_bits** p : {&c->phi16[0], &c->phi16[1], &c->phit16, &c->phia16, &c->phiT16, &c->phitT16, &c->phiaT16, &c->gt16,
                        &c->h16, &c->nsq16})
    TRY(dalloc(c, *p, kk2));
  if (c->have_y && !c->y16_valid) {
    TRY(f32_to_bf16(c->s, (int64_t)ny, c->YA, c->YA16));
    TRY(f32_to_bf16(c->s, (int64_t)ny, c->YB, c->YB16));
    TRY(f32_to_bf16(c->s, (int64_t)ny, c->YCt, c->YCt16));
    c->y16_valid = true;
  }
  return MMSB_OK;
}

// Grows the ||dPhi||^2 log to hold `more` further steps.
static int ensure_frob(mmsb_ctx* c, int64_t more) {
  if (c->steps_done + more <= c->frob_cap) return MMSB_OK;
  int64_t cap = std::max<int64_t>(1024, 2 * (c->steps_done + more));
  double* nf = nullptr;
  HCHK(hipMalloc(&nf, sizeof(double) * cap));
  HCHK(hipMemsetAsync(nf, 0, sizeof(double) * cap, c->s));
  if (c->frob && c->steps_done)
    HCHK(hipMemcpyAsync(nf, c->frob, sizeof(double) * c->steps_done, hipMemcpyDeviceToDevice, c->s));
  if (c->frob) {
    HCHK(hipStreamSynchronize(c->s));
    auto it = std::find(c->allocs.begin(), c->allocs.end(), (void*)c->frob);
    if (it != c->allocs.end()) c->allocs.erase(it);
    HCHK(hipFree(c->frob));
  }
  c->allocs.push_back(nf);
  c->frob = nf;
  c->frob_cap = cap;
  return MMSB_OK;
}

static StgdArgs stgd_args(const mmsb_ctx* c) {
  StgdArgs a{};
  a.k = c->k;
  a.kp = c->kp;
  a.b = c->b;
  a.bl = c->bl;
  a.n_x = c->n[0];
  a.n_pad = c->n_pad;
  // the T of the same reference path as the M2 that whitened the vectors: paper M2 with the paper
  // coefficients (CPU Util.cpp:289-298), GPU M2 with init_buff's (stpm.c:1026-1038)
  double co[3];
  (void)mmsb_tensor_coeffs(c->cfg.alpha_0, c->cfg.m2_scaling, co);
  a.c3 = (float)co[0];
  a.c1 = (float)co[1];
  a.c0 = (float)co[2];
  a.eta = (float)c->cfg.learn_rate;
  a.theta = (float)c->cfg.theta;
  a.mu_a = c->muw[0];
  a.mu_b = c->muw[1];
  a.mu_c = c->muw[2];
  a.cs_part = c->cs_part;
  a.mab = c->mab;
  a.frob_pre = c->cfg.renorm_mode == MMSB_RENORM_QR;  // the reference logs before its QR
  return a;
}

static StgdBuffers stgd_bufs(const mmsb_ctx* c) {
  StgdBuffers bufs{};
  bufs.phi[0] = c->phi[0];
  bufs.phi[1] = c->phi[1];
  bufs.phit = c->phit;
  bufs.phia = c->phia;
  bufs.gt = c->gt;
  bufs.phiT = c->phiT;
  bufs.phitT = c->phitT;
  bufs.phiaT = c->phiaT;
  bufs.h = c->h;
  bufs.q = c->nsq;
  bufs.phi16[0] = c->phi16[0];
  bufs.phi16[1] = c->phi16[1];
  bufs.phit16 = c->phit16;
  bufs.phia16 = c->phia16;
  bufs.phiT16 = c->phiT16;
  bufs.phitT16 = c->phitT16;
  bufs.phiaT16 = c->phiaT16;
  bufs.gt16 = c->gt16;
  bufs.h16 = c->h16;
  bufs.q16 = c->nsq16;
  bufs.cgt = c->cgt;
  bufs.rowsum = c->ns_rowsum;
  bufs.mpart = c->mpart;
  bufs.fpart = c->fpart;
  return bufs;
}

static void y_ptrs(const mmsb_ctx* c, int32_t precision, const void*& ya, const void*& yb, const void*& yc) {
  const bool bf = precision == MMSB_PREC_BF16;
  ya = bf ? (const void*)c->YA16 : (const void*)c->YA;
  yb = bf ? (const void*)c->YB16 : (const void*)c->YB;
  yc = bf ? (const void*)c->YCt16 : (const void*)c->YCt;
}

// The end test of the reference loop (stpm.c:747-765) after global step t with log value f.
static bool stop_rule(const stpm_config& cf, int64_t t, double f) {
  return (cf.sgd_iter_min > 0 && f < cf.thresh && t > cf.sgd_iter_min) || (cf.sgd_iter_max > 0 && t > cf.sgd_iter_max);
}

// Runs nsteps steps. With `rule`, the end test of the reference loop is applied after every step
// (on the device within a launch, on the host at the end of each launch) and the run stops after the
// first step that meets it; *stopped reports that.
static int run_stgd_steps(mmsb_ctx* c, int64_t nsteps, int32_t precision, StgdTiming* timing = nullptr,
                          bool rule = false, bool* stopped = nullptr) {
  if (stopped) *stopped = false;
  if (nsteps <= 0) return MMSB_OK;
  if (precision != MMSB_PREC_FP32 && precision != MMSB_PREC_BF16)
    return set_error(MMSB_ERR_ARG, "stgd: unknown precision %d", precision);
  if (precision == MMSB_PREC_BF16) TRY(prepare_bf16(c));
  TRY(ensure_frob(c, nsteps));
  const bool qr = c->cfg.renorm_mode == MMSB_RENORM_QR;
  if (qr) TRY(alloc_small(c));
  StgdArgs a = stgd_args(c);
  if (rule) {
    TRY(dalloc(c, c->stopctl, 1));
    a.stop = c->stopctl;
    a.thresh = c->cfg.thresh;
    a.iter_min = c->cfg.sgd_iter_min;
    a.iter_max = c->cfg.sgd_iter_max;
  }
  const StgdBuffers bufs = stgd_bufs(c);
  const void *ya, *yb, *yc;
  y_ptrs(c, precision, ya, yb, yc);
  StgdFill fill;
  if (!c->have_y) {  // checkpointed: the minibatch's rows are rebuilt from CSR before each step
    a.y_window = 1;
    a.n_pad = c->b;
    const bool bf = precision == MMSB_PREC_BF16;
    ya = bf ? (const void*)c->yw16[0] : (const void*)c->yw[0];
    yb = bf ? (const void*)c->yw16[1] : (const void*)c->yw[1];
    yc = bf ? (const void*)c->yw16[2] : (const void*)c->yw[2];
    fill = [c, bf](int64_t row0, int32_t nb) { return fill_window(c, bf, row0, nb); };
  }
  StgdRenorm renorm;
  if (qr)  // CholeskyQR2 of Phi~ (k x k, FP64 Gram and Cholesky): R with a positive diagonal, as Dgeqrfp
    renorm = [c](const float* src, float* dst) {
      return orth(c, const_cast<float*>(src), c->k, c->k, c->kp, dst, c->nsq, nullptr, 0);
    };
  const int64_t spe = (c->n[0] + c->bl - 1) / c->bl;
  auto launch = [&](int64_t step0, int64_t ns, double* frob_out, int& cur, StgdTiming* tm) -> int {
    return stgd_run(c->s, precision, a, ya, yb, yc, step0, ns, c->cfg.renorm_every, c->cfg.ns_iters, c->cfg.ns_degree, bufs,
                    cur, frob_out, tm, fill, renorm);
  };
  // Whole epochs that start at an epoch boundary replay a captured HIP graph of the epoch's
  // launches (one graph per precision, Phi ping-pong parity and end-test setting); anything else
  // launches directly.
  const bool renorm_aligned = c->cfg.renorm_every <= 1 || spe % c->cfg.renorm_every == 0;
  if (rule) TRY(stgd_set_i64(c->s, &c->stopctl->stop_at, INT64_MAX));
  int64_t done = 0;
  while (done < nsteps) {
    const int64_t step0 = c->steps_done + done;
    const int cur0 = c->cur;
    int64_t ns = 0;
    if (rule) TRY(stgd_set_i64(c->s, &c->stopctl->base, step0));
    const bool epoch_aligned =
        (step0 % spe == 0) && (nsteps - done >= spe) && renorm_aligned && c->cfg.use_graphs && !timing;
    if (epoch_aligned) {
      TRY(dalloc(c, c->frob_epoch, (size_t)spe));
      hipGraphExec_t& ge = c->epoch_graph[precision][c->cur][rule ? 1 : 0];
      if (!ge) {
        hipGraph_t g = nullptr;
        int cur = c->cur;
        HCHK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
        const int st = launch(0, spe, c->frob_epoch, cur, nullptr);
        hipError_t e = hipStreamEndCapture(c->s, &g);
        if (st != MMSB_OK) {
          if (g) (void)hipGraphDestroy(g);
          return st;
        }
        HCHK(e);
        HCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
        HCHK(hipGraphDestroy(g));
      }
      HCHK(hipGraphLaunch(ge, c->s));
      HCHK(hipMemcpyAsync(c->frob + step0, c->frob_epoch, sizeof(double) * spe, hipMemcpyDeviceToDevice, c->s));
      if (spe & 1) c->cur ^= 1;
      ns = spe;
    } else {
      ns = std::min<int64_t>(nsteps - done, spe - step0 % spe);
      TRY(launch(step0, ns, c->frob + step0, c->cur, timing));
    }
    if (rule) {  // where did the launch stop, and does its last step end the loop?
      StgdStop hs{};
      HCHK(hipMemcpyAsync(&hs, c->stopctl, sizeof(hs), hipMemcpyDeviceToHost, c->s));
      HCHK(hipStreamSynchronize(c->s));
      int64_t ran = ns;
      if (hs.stop_at < step0 + ns) {
        ran = std::max<int64_t>(0, hs.stop_at - step0);
        c->cur = cur0 ^ (int)(ran & 1);
      }
      done += ran;
      if (ran < ns) {
        if (stopped) *stopped = true;
        break;
      }
      double last = 0.0;
      HCHK(hipMemcpyAsync(&last, c->frob + step0 + ns - 1, sizeof(double), hipMemcpyDeviceToHost, c->s));
      HCHK(hipStreamSynchronize(c->s));
      if (stop_rule(c->cfg, step0 + ns - 1, last)) {
        if (stopped) *stopped = true;
        break;
      }
    } else {
      done += ns;
    }
    if (done < nsteps) {  // a stop request ends the run at this epoch boundary
      HCHK(hipStreamSynchronize(c->s));
      if (c->stop_req.load()) break;
    }
  }
  c->steps_done += done;
  if (qr) TRY(check_status(c, "QR renormalisation"));
  return MMSB_OK;
}

}  // namespace mmsb

// ---- many instances: one persistent launch per epoch -------------------------------------------
// The plan of each epoch parity covers one epoch of every member (members may have different n, k
// and minibatch) and runs as one cooperative launch. A member's Phi ping-pong flips every epoch
// when its steps per epoch are odd, so parity 1 differs from parity 0 exactly in those members' cur.
struct mmsb_batch {
  std::vector<mmsb_ctx*> ctx;
  int32_t precision = MMSB_PREC_FP32;
  hipStream_t s = nullptr;
  int device = 0;
  bool slots = false;          // true: grouped per-slot launches replayed from a graph per parity
  bool coop_refused = false;   // a cooperative launch was refused: grouped launches from then on
  StgdBatch* plan[2] = {nullptr, nullptr};
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  std::vector<int> cur0;       // members' cur at even batch epochs
  std::vector<int64_t> spe;    // steps per epoch
  int64_t epochs_done = 0;
  std::vector<float*> save;    // autotuning: snapshots of each member's Phi
};

static int64_t steps_per_epoch(const mmsb_ctx* c) { return (c->n[0] + c->bl - 1) / c->bl; }

static void batch_drop_graphs(mmsb_batch* b) {
  for (hipGraphExec_t& g : b->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// One epoch of plan p, by the batch's current mode. A refused cooperative launch (another process
// on the GPU can hold the global-wave-sync resource it needs) enqueues nothing, so the epoch then
// runs as grouped launches, and so do the later ones.
static int batch_epoch(mmsb_batch* b, int p) {
  if (!b->slots) {
    if (stgd_batch_launch(b->s, b->plan[p]) == MMSB_OK) return MMSB_OK;
    (void)hipGetLastError();
    clear_error();
    b->coop_refused = true;
    b->slots = true;
  }
  if (!b->graph[p]) {
    hipGraph_t g = nullptr;
    HCHK(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
    const int st = stgd_batch_launch_slots(b->s, b->plan[p]);
    hipError_t ce = hipStreamEndCapture(b->s, &g);
    if (st != MMSB_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    HCHK(ce);
    hipError_t ie = hipGraphInstantiate(&b->graph[p], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HCHK(ie);
  }
  HCHK(hipGraphLaunch(b->graph[p], b->s));
  return MMSB_OK;
}

extern "C" void mmsb_batch_destroy(mmsb_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->s) (void)hipStreamSynchronize(b->s);
  batch_drop_graphs(b);
  for (StgdBatch*& p : b->plan) stgd_batch_free(p);
  for (float* p : b->save) (void)hipFree(p);
  delete b;
}

extern "C" int mmsb_batch_create(mmsb_ctx* const* ctxs, int32_t count, int32_t precision, mmsb_batch** out) {
  clear_error();
  MMSB_CHECK_ARG(ctxs && out && count >= 1, "batch_create: bad arguments");
  *out = nullptr;
  MMSB_CHECK_ARG(precision == MMSB_PREC_FP32 || precision == MMSB_PREC_BF16, "batch_create: unknown precision %d", precision);
  const mmsb_ctx* c0 = ctxs[0];
  MMSB_CHECK_ARG(c0, "batch_create: null context 0");
  for (int32_t i = 0; i < count; ++i) {
    const mmsb_ctx* c = ctxs[i];
    MMSB_CHECK_ARG(c, "batch_create: null context %d", i);
    if (!c->have_y)
      return set_error(MMSB_ERR_STATE, c->y_window ? "batch_create: context %d recomputes its minibatches (y_mode); "
                                                     "batches need resident neighbour vectors"
                                                   : "batch_create: context %d has no neighbour vectors", i);
    MMSB_CHECK_ARG(c->cfg.device == c0->cfg.device, "batch_create: contexts on different devices");
    MMSB_CHECK_ARG(c->cfg.ns_iters == c0->cfg.ns_iters && c->cfg.ns_degree == c0->cfg.ns_degree &&
                       c->cfg.renorm_every == c0->cfg.renorm_every,
                   "batch_create: contexts disagree on ns_iters / ns_degree / renorm_every");
    const int64_t spe = steps_per_epoch(c);
    MMSB_CHECK_ARG(c->steps_done % spe == 0, "batch_create: context %d is not at an epoch boundary", i);
    MMSB_CHECK_ARG(c->cfg.renorm_every <= 1 || spe % c->cfg.renorm_every == 0,
                   "batch_create: context %d: renorm_every must divide the steps per epoch", i);
    for (int32_t j = 0; j < i; ++j) MMSB_CHECK_ARG(ctxs[j] != c, "batch_create: context %d appears twice", i);
  }
  HCHK(hipSetDevice(c0->cfg.device));
  auto* b = new mmsb_batch();
  b->precision = precision;
  b->device = c0->cfg.device;
  b->s = c0->s;
  auto fail = [&](int st) {
    mmsb_batch_destroy(b);
    return st;
  };
  std::vector<StgdBatchMember> mem[2];
  for (int32_t i = 0; i < count; ++i) {
    mmsb_ctx* c = ctxs[i];
    int st = alloc_stgd(c);
    if (st == MMSB_OK && !c->have_phi) st = init_phi(c);
    if (st == MMSB_OK && precision == MMSB_PREC_BF16) st = prepare_bf16(c);
    const int64_t spe = steps_per_epoch(c);
    if (st == MMSB_OK) st = dalloc(c, c->frob_epoch, (size_t)spe);
    if (st != MMSB_OK) return fail(st);
    if (hipStreamSynchronize(c->s) != hipSuccess) return fail(set_error(MMSB_ERR_HIP, "batch_create: stream sync failed"));
    b->ctx.push_back(c);
    b->cur0.push_back(c->cur);
    b->spe.push_back(spe);
    for (int p = 0; p < 2; ++p) {
      StgdBatchMember m{};
      m.args = stgd_args(c);
      m.bufs = stgd_bufs(c);
      y_ptrs(c, precision, m.YA, m.YB, m.YCt);
      m.cur = c->cur ^ (int)((p * spe) & 1);
      m.step0 = 0;
      m.nsteps = spe;
      m.frob_log = c->frob_epoch;
      mem[p].push_back(m);
    }
  }
  for (int p = 0; p < 2; ++p) {
    const int st = stgd_batch_plan(b->plan[p], precision, mem[p], c0->cfg.ns_iters, c0->cfg.ns_degree, c0->cfg.renorm_every);
    if (st != MMSB_OK) return fail(st);
  }
  *out = b;
  return MMSB_OK;
}

static int batch_check_state(mmsb_batch* b) {
  const int parity = (int)(b->epochs_done & 1);
  for (size_t i = 0; i < b->ctx.size(); ++i) {
    const mmsb_ctx* c = b->ctx[i];
    if (c->steps_done % b->spe[i] != 0 || c->cur != (b->cur0[i] ^ (int)((parity * b->spe[i]) & 1)))
      return set_error(MMSB_ERR_STATE, "batch: member %zu advanced outside the batch; create a new batch", i);
  }
  return MMSB_OK;
}

extern "C" int mmsb_batch_run(mmsb_batch* b, int32_t epochs) {
  clear_error();
  MMSB_CHECK_ARG(b && epochs >= 0, "batch_run: bad arguments");
  HCHK(hipSetDevice(b->device));
  TRY(batch_check_state(b));
  for (mmsb_ctx* c : b->ctx) {
    TRY(ensure_frob(c, (int64_t)epochs * steps_per_epoch(c)));
    HCHK(hipStreamSynchronize(c->s));
  }
  for (int32_t e = 0; e < epochs; ++e) {
    const int p = (int)((b->epochs_done + e) & 1);
    TRY(batch_epoch(b, p));
    for (size_t i = 0; i < b->ctx.size(); ++i) {
      mmsb_ctx* c = b->ctx[i];
      HCHK(hipMemcpyAsync(c->frob + c->steps_done, c->frob_epoch, sizeof(double) * b->spe[i], hipMemcpyDeviceToDevice, b->s));
      c->steps_done += b->spe[i];
      if (b->spe[i] & 1) c->cur ^= 1;
      c->eig_done = c->post_done = false;
    }
  }
  b->epochs_done += epochs;
  HCHK(hipStreamSynchronize(b->s));
  return MMSB_OK;
}

extern "C" int mmsb_batch_autotune(mmsb_batch* b, int32_t* best_grid, double* best_ms) {
  clear_error();
  MMSB_CHECK_ARG(b, "batch_autotune: null batch");
  HCHK(hipSetDevice(b->device));
  TRY(batch_check_state(b));
  for (mmsb_ctx* c : b->ctx) HCHK(hipStreamSynchronize(c->s));
  StgdBatch* plan = b->plan[b->epochs_done & 1];
  if (b->save.empty())
    for (mmsb_ctx* c : b->ctx) {
      float* q = nullptr;
      HCHK(hipMalloc(&q, sizeof(float) * (size_t)c->kp * c->kp));
      b->save.push_back(q);
    }
  for (size_t i = 0; i < b->ctx.size(); ++i) {
    const mmsb_ctx* c = b->ctx[i];
    HCHK(hipMemcpyAsync(b->save[i], c->phi[c->cur], sizeof(float) * (size_t)c->kp * c->kp, hipMemcpyDeviceToDevice, b->s));
  }
  auto restore = [&]() -> int {
    for (size_t i = 0; i < b->ctx.size(); ++i) {
      const mmsb_ctx* c = b->ctx[i];
      HCHK(hipMemcpyAsync(c->phi[c->cur], b->save[i], sizeof(float) * (size_t)c->kp * c->kp, hipMemcpyDeviceToDevice,
                          b->s));
    }
    return MMSB_OK;
  };
  int dev = 0, cus = 0;
  HCHK(hipGetDevice(&dev));
  HCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int32_t gmax = stgd_batch_grid_max(plan);
  std::vector<int32_t> cands;  // persistent grid sizes; 0 = grouped per-slot launches from a graph
  for (int32_t g : {cus / 4, cus / 2, cus, 2 * cus, 3 * cus, 4 * cus})
    if (g >= 1 && g <= gmax && std::find(cands.begin(), cands.end(), g) == cands.end()) cands.push_back(g);
  cands.push_back(0);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HCHK(hipEventCreate(&e0));
  HCHK(hipEventCreate(&e1));
  const int32_t keep = stgd_batch_grid(plan);
  const bool keep_slots = b->slots;
  int32_t best = keep_slots ? 0 : keep;
  double bms = 1e300;
  int st = MMSB_OK;
  const int p = (int)(b->epochs_done & 1);
  for (int32_t g : cands) {
    if (g > 0 && b->coop_refused) continue;  // persistent candidates cannot launch here
    b->slots = g == 0;
    if (g > 0) stgd_batch_set_grid(plan, g);
    for (int rep = 0; rep < 2 && st == MMSB_OK; ++rep) {
      if ((st = restore()) != MMSB_OK) break;
      float ms = 0.f;
      if (hipEventRecord(e0, b->s) != hipSuccess) st = set_error(MMSB_ERR_HIP, "batch_autotune: event record failed");
      if (st == MMSB_OK) st = batch_epoch(b, p);
      if (st == MMSB_OK && hipEventRecord(e1, b->s) != hipSuccess) st = set_error(MMSB_ERR_HIP, "batch_autotune: event record failed");
      if (st == MMSB_OK && hipEventSynchronize(e1) != hipSuccess) st = set_error(MMSB_ERR_HIP, "batch_autotune: sync failed");
      if (st == MMSB_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = set_error(MMSB_ERR_HIP, "batch_autotune: timing failed");
      if (g > 0 && b->coop_refused) break;        // timed the grouped fallback, not this candidate
      if (st == MMSB_OK && rep == 1 && ms < bms) {  // the first repetition of a mode may include graph set-up
        bms = ms;
        best = g;
      }
    }
    if (st != MMSB_OK) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != MMSB_OK) {
    b->slots = keep_slots;
    stgd_batch_set_grid(plan, keep);
  }
  TRY(st);
  TRY(restore());
  HCHK(hipStreamSynchronize(b->s));
  if (b->coop_refused) best = 0;
  b->slots = best == 0;
  for (StgdBatch* q : b->plan)
    if (best > 0) stgd_batch_set_grid(q, best);
  if (best_grid) *best_grid = best;
  if (best_ms) *best_ms = bms;
  return MMSB_OK;
}

namespace mmsb {

// ---- eigenvalues and Pi ------------------------------------------------------------------------
static int eigen_impl(mmsb_ctx* c) {
  const int32_t k = c->k;
  TRY(dalloc(c, c->lam, k));
  TRY(dalloc(c, c->lamp, k));
  TRY(dalloc(c, c->ahat, k));
  TRY(dalloc(c, c->phil, (size_t)k * k));
  TRY(eig_post(c->s, k, c->kp, c->phi[c->cur], c->cfg.pinv_tol, c->lam, c->lamp, c->ahat, c->phil));
  c->eig_done = true;
  return MMSB_OK;
}

static int post_impl(mmsb_ctx* c) {
  hipStream_t s = c->s;
  const int32_t k = c->k, kp = c->kp;
  if (!c->whitened) return set_error(MMSB_ERR_STATE, "post_proc needs the whitening of this context");
  if (!c->eig_done) TRY(eigen_impl(c));
  TRY(dalloc(c, c->Upost, (size_t)c->n[1] * kp));
  TRY((gemm_nn<float, float>(s, c->n[1], k, k, 1.0, c->W, kp, c->phil, k, 0.0, c->Upost, kp)));  // W Phi Lambda'
  TRY(dalloc(c, c->Pi[0], (size_t)c->n[0] * kp));
  TRY(spmm(c, c->blk[MMSB_BLOCK_XA], 1.f, c->Upost, kp, k, c->Pi[0], kp));                           // post_proc.c:93
  if (c->have[MMSB_BLOCK_BA]) {
    TRY(dalloc(c, c->LBU, (size_t)k * k));
    TRY(dalloc(c, c->ZBtU, (size_t)c->n[2] * kp));
    TRY((gemm_tn<float, float>(s, c->n[1], k, k, 1.0, c->LB, kp, c->Upost, kp, 0.0, c->LBU, c->work)));
    TRY((gemm_nn<float, float>(s, c->n[2], k, k, 1.0, c->U, kp, c->LBU, k, 0.0, c->ZBtU, kp)));      // Z_B^T Upost
    TRY(dalloc(c, c->Pi[1], (size_t)c->n[1] * kp));
    TRY(spmm(c, c->blkT[MMSB_BLOCK_BA], 1.f, c->ZBtU, kp, k, c->Pi[1], kp));                         // post_proc.c:97-100
    TRY(dalloc(c, c->Pi[2], (size_t)c->n[2] * kp));
    TRY(spmm(c, c->blk[MMSB_BLOCK_BA], 1.f, c->Upost, kp, k, c->Pi[2], kp));                          // post_proc.c:103
  }
  if (c->have[MMSB_BLOCK_CA]) {
    TRY(dalloc(c, c->Pi[3], (size_t)c->n[3] * kp));
    TRY(spmm(c, c->blk[MMSB_BLOCK_CA], 1.f, c->Upost, kp, k, c->Pi[3], kp));                          // post_proc.c:106
  }
  const int64_t rows[4] = {c->n[0], c->n[1], c->n[2], c->n[3]};
  for (int p = 0; p < 4; ++p) {
    if (!c->Pi[p]) continue;
    TRY(dalloc(c, c->PiN[p], (size_t)rows[p] * kp));
    TRY(dalloc(c, c->cnt[p], (size_t)rows[p]));
    TRY(norm_prob(s, rows[p], k, c->Pi[p], kp, c->PiN[p], kp, (float)c->cfg.member_tau, c->cnt[p]));
  }
  c->post_done = true;
  return MMSB_OK;
}

// ---- getters -------------------------------------------------------------------------------------
struct View {
  const void* p = nullptr;
  int dtype = 0;  // 0 float, 1 double, 2 int32
  int64_t rows = 0;
  int32_t cols = 0;
  int64_t ld = 0;
  bool trans = false;  // stored [cols x rows] with leading dimension ld
};

static View view_of(mmsb_ctx* c, int what) {
  View v;
  const int32_t k = c->k, kp = c->kp;
  auto tall = [&](const void* p, int64_t rows) { View w; w.p = p; w.rows = rows; w.cols = k; w.ld = kp; return w; };
  auto vec = [&](const void* p, int64_t n) { View w; w.p = p; w.dtype = 1; w.rows = n; w.cols = 1; w.ld = 1; return w; };
  switch (what) {
    case MMSB_Q_MU_A: case MMSB_Q_MU_B: case MMSB_Q_MU_C:
      return c->mu[what - MMSB_Q_MU_A] ? vec(c->mu[what - MMSB_Q_MU_A], c->n[1 + what - MMSB_Q_MU_A]) : v;
    case MMSB_Q_W: return c->whitened ? tall(c->W, c->n[1]) : v;
    case MMSB_Q_U: return c->whitened ? tall(c->U, c->n[2]) : v;
    case MMSB_Q_V: return c->whitened && c->V ? tall(c->V, c->n[3]) : v;
    case MMSB_Q_SIGMA: return c->whitened ? vec(c->sig, k) : v;
    case MMSB_Q_LB: return c->whitened ? tall(c->LB, c->n[1]) : v;
    case MMSB_Q_LC: return c->whitened ? tall(c->LC, c->n[1]) : v;
    case MMSB_Q_LAM_M2: return c->whitened ? vec(c->lamv, k) : v;
    case MMSB_Q_YA: return c->have_y ? tall(c->YA, c->n[0]) : v;
    case MMSB_Q_YB: return c->have_y ? tall(c->YB, c->n[0]) : v;
    case MMSB_Q_YC:
      if (!c->have_y) return v;
      v.p = c->YCt; v.rows = c->n[0]; v.cols = k; v.ld = c->n_pad; v.trans = true;
      return v;
    case MMSB_Q_MUW_A: case MMSB_Q_MUW_B: case MMSB_Q_MUW_C:
      return c->have_y || (c->y_window && c->whitened) ? vec(c->muw[what - MMSB_Q_MUW_A], k) : v;
    case MMSB_Q_PHI:
      if (!c->have_phi) return v;
      v.p = c->phi[c->cur]; v.rows = k; v.cols = k; v.ld = kp;
      return v;
    case MMSB_Q_FROB: return c->steps_done ? vec(c->frob, c->steps_done) : v;
    case MMSB_Q_LAMBDA: return c->eig_done ? vec(c->lam, k) : v;
    case MMSB_Q_LAMBDA_P: return c->eig_done ? vec(c->lamp, k) : v;
    case MMSB_Q_ALPHA_HAT: return c->eig_done ? vec(c->ahat, k) : v;
    case MMSB_Q_PI_X: case MMSB_Q_PI_A: case MMSB_Q_PI_B: case MMSB_Q_PI_C: {
      const int p = what - MMSB_Q_PI_X;
      return c->Pi[p] ? tall(c->Pi[p], c->n[p]) : v;
    }
    case MMSB_Q_PIN_X: case MMSB_Q_PIN_A: case MMSB_Q_PIN_B: case MMSB_Q_PIN_C: {
      const int p = what - MMSB_Q_PIN_X;
      return c->PiN[p] ? tall(c->PiN[p], c->n[p]) : v;
    }
    case MMSB_Q_COUNT_X: case MMSB_Q_COUNT_A: case MMSB_Q_COUNT_B: case MMSB_Q_COUNT_C: {
      const int p = what - MMSB_Q_COUNT_X;
      if (!c->cnt[p]) return v;
      v.p = c->cnt[p]; v.dtype = 2; v.rows = c->n[p]; v.cols = 1; v.ld = 1;
      return v;
    }
    default: return v;
  }
}

// Device -> host copy of a (possibly strided or transposed) quantity: compacted on the device into a
// contiguous staging buffer, then linear copies (strided 2D copies of ~1e6 short rows into pageable
// memory are orders of magnitude slower). The staging goes in slabs of at most 256 MiB, so that a
// copy of a tall matrix does not need a second full-size device allocation.
static int copy_raw(mmsb_ctx* c, const View& v, void* host) {
  const size_t esz = v.dtype == 1 ? 8 : 4;
  const int64_t outer = v.trans ? v.cols : v.rows;
  const int64_t inner = v.trans ? v.rows : v.cols;
  const size_t bytes = (size_t)outer * inner * esz;
  if (v.ld == inner) return d2h(c->s, host, v.p, bytes);
  const int64_t slab = std::max<int64_t>(1, std::min<int64_t>(outer, ((int64_t)256 << 20) / std::max<int64_t>(1, inner * esz)));
  void* d = nullptr;
  HCHK(hipMalloc(&d, std::max<size_t>((size_t)slab * inner * esz, 1)));
  int st = MMSB_OK;
  for (int64_t o = 0; o < outer && st == MMSB_OK; o += slab) {
    const int64_t m = std::min<int64_t>(slab, outer - o);
    hipError_t e = hipMemcpy2DAsync(d, inner * esz, static_cast<const unsigned char*>(v.p) + (size_t)o * v.ld * esz,
                                    v.ld * esz, inner * esz, m, hipMemcpyDeviceToDevice, c->s);
    if (e != hipSuccess) st = set_error(MMSB_ERR_HIP, "copy_out: %s", hipGetErrorString(e));
    if (st == MMSB_OK) st = d2h(c->s, static_cast<unsigned char*>(host) + (size_t)o * inner * esz, d, (size_t)m * inner * esz);
  }
  (void)hipFree(d);
  return st;
}

static int copy_out(mmsb_ctx* c, const View& v, double* host) {
  const size_t esz = v.dtype == 1 ? 8 : 4;
  std::vector<unsigned char> buf((size_t)v.rows * v.cols * esz);
  TRY(copy_raw(c, v, buf.data()));
  const int64_t total = v.rows * v.cols;
  parallel_for(total, default_threads(), [&](int64_t lo, int64_t hi) {
    for (int64_t t = lo; t < hi; ++t) {
      const int64_t i = t / v.cols, j = t % v.cols;
      const int64_t src = v.trans ? j * v.rows + i : t;
      double x;
      if (v.dtype == 0) x = reinterpret_cast<const float*>(buf.data())[src];
      else if (v.dtype == 1) x = reinterpret_cast<const double*>(buf.data())[src];
      else x = reinterpret_cast<const int32_t*>(buf.data())[src];
      host[t] = x;
    }
  });
  return MMSB_OK;
}

static std::string out_name(const stpm_config& cf, const char* stem) {
  return std::string(cf.folder_name) + stem + "_" + std::to_string(cf.ncom) + "_" + cf.alpha_0_str + "_" + cf.learn_rate_str + ".txt";
}

// The CPU program's two result files (CommunityDetection.cpp:121-144, write_pi IOfn.cpp:57-67):
// pi_path gets the k x (n_X + n_A + n_B + n_C) matrix of normalised memberships (the partitions
// concatenated in the order X, A, B, C, each column a node's normProbVector) and white_path the
// n_A x k whitening matrix W, both as "row+1 \t col+1 \t value" triplets of the nonzeros in
// column-major order with 6 significant digits. A null path skips that file. The normalised Pi is
// fetched in slabs of rows, so the host holds one slab, not the whole matrix.
static int write_triplets_impl(mmsb_ctx* c, const char* pi_path, const char* white_path) {
  const int32_t k = c->k;
  if (pi_path) {
    if (!c->post_done) return set_error(MMSB_ERR_STATE, "write_triplets: Pi not computed (call mmsb_post)");
    FILE* f = std::fopen(pi_path, "w");
    if (!f) return set_error(MMSB_ERR_IO, "writing %s failed", pi_path);
    int64_t col0 = 0;
    int st = MMSB_OK;
    for (int p = 0; p < 4 && st == MMSB_OK; ++p) {
      View v = view_of(c, MMSB_Q_PIN_X + p);
      if (!v.p) continue;
      const int64_t rows = v.rows, slab = std::max<int64_t>(1, ((int64_t)64 << 20) / ((int64_t)k * 8));
      std::vector<double> h;
      for (int64_t r0 = 0; r0 < rows && st == MMSB_OK; r0 += slab) {
        View w = v;
        w.rows = std::min(slab, rows - r0);
        w.p = static_cast<const float*>(v.p) + r0 * v.ld;
        h.resize((size_t)w.rows * k);
        st = copy_out(c, w, h.data());
        for (int64_t i = 0; i < w.rows && st == MMSB_OK; ++i)
          for (int32_t j = 0; j < k; ++j) {
            const double x = h[(size_t)i * k + j];
            if (x != 0.0) std::fprintf(f, "%d\t%lld\t%g\n", j + 1, (long long)(col0 + r0 + i + 1), x);
          }
      }
      col0 += rows;
    }
    if (std::fclose(f) != 0 && st == MMSB_OK) st = set_error(MMSB_ERR_IO, "closing %s failed", pi_path);
    TRY(st);
  }
  if (white_path) {
    if (!c->whitened) return set_error(MMSB_ERR_STATE, "write_triplets: W not computed (call mmsb_whiten)");
    const View v = view_of(c, MMSB_Q_W);
    std::vector<float> h((size_t)v.rows * k);
    TRY(copy_raw(c, v, h.data()));
    FILE* f = std::fopen(white_path, "w");
    if (!f) return set_error(MMSB_ERR_IO, "writing %s failed", white_path);
    for (int32_t j = 0; j < k; ++j)
      for (int64_t a = 0; a < v.rows; ++a) {
        const double x = h[(size_t)a * k + j];
        if (x != 0.0) std::fprintf(f, "%lld\t%d\t%g\n", (long long)a + 1, j + 1, x);
      }
    if (std::fclose(f) != 0) return set_error(MMSB_ERR_IO, "closing %s failed", white_path);
  }
  return MMSB_OK;
}

static int set_sizes(mmsb_ctx* c) {
  const stpm_config& cf = c->cfg;
  c->k = cf.ncom;
  c->r = cf.ncom + std::max(0, cf.oversample);
  c->kp = (int32_t)r64(c->k, 64);
  c->rld = (int32_t)r64(c->r, 4);
  c->rmax = std::max(c->r, c->kp);
  c->bl = std::max(1, cf.minibatch);
  c->b = (int32_t)r64(c->bl, 64);
  c->n[0] = cf.nx;
  c->n[1] = cf.na;
  c->n[2] = cf.nb;
  c->n[3] = cf.nc;
  c->nmax = std::max(std::max(c->n[0], c->n[1]), std::max(c->n[2], c->n[3]));
  c->n_pad = r64((r64(c->n[0], c->bl) - c->bl + c->b), 64);
  return MMSB_OK;
}

}  // namespace mmsb

// ===================================== C-ABI =========================================================
extern "C" size_t stpm_config_size(void) { return sizeof(stpm_config); }

extern "C" void stpm_config_default(stpm_config* c) {
  std::memset(c, 0, sizeof(*c));
  c->nx = 29079;  // stpm.h:72-76
  c->na = 29080;
  c->nb = 29079;
  c->nc = 29079;
  c->ncom = 100;
  c->alpha_0 = 1.0;       // stpm.h:78
  c->learn_rate = 1e-3;   // stpm.h:79
  c->thresh = 1.0;        // stpm.h:94
  c->pinv_tol = 1e-6;     // stpm.h:95
  c->sgd_iter_min = 0;    // 0: run every configured epoch
  c->sgd_iter_max = 0;
  c->binary = 1;
  c->synthetic = 1;
  std::snprintf(c->folder_name, sizeof(c->folder_name), "DBLP_116317/");
  std::snprintf(c->file_prefix, sizeof(c->file_prefix), "DBLP_");
  std::snprintf(c->alpha_0_str, sizeof(c->alpha_0_str), "1.0");
  std::snprintf(c->learn_rate_str, sizeof(c->learn_rate_str), "1e-3");
  c->minibatch = 1024;
  c->epochs = 1;
  c->theta = 1.0;
  c->renorm_every = 1;
  c->ns_iters = 2;
  c->ns_degree = 3;
  c->precision = MMSB_PREC_FP32;
  c->polish_epochs = 1;
  c->oversample = 10;
  c->power_iters = 1;
  c->m2_scaling = MMSB_M2_PAPER;
  c->member_tau = 0.0;
  c->seed = 0;
  c->ba_first_is_row = 1;
  c->feof_repeat = 0;
  c->device = 0;
  c->verbose = 0;
  c->use_graphs = 1;
}

// Live contexts: the process-wide staging and sort temporaries are released with the last one.
static std::atomic<int> g_live_contexts{0};

static void release_process_resources() {
  release_staging();
  release_streamer();
  release_sort_scratch();
}

extern "C" int mmsb_create(const stpm_config* cfg, mmsb_ctx** out) {
  clear_error();
  MMSB_CHECK_ARG(cfg && out, "create: null argument");
  MMSB_CHECK_ARG(cfg->nx > 0 && cfg->na > 0 && cfg->nb > 0 && cfg->nc > 0, "create: partition sizes must be positive");
  MMSB_CHECK_ARG(cfg->ncom >= 1 && cfg->ncom <= 2048, "create: NCOM=%d outside [1, 2048]", cfg->ncom);
  MMSB_CHECK_ARG(cfg->ncom + cfg->oversample <= 4096 && cfg->oversample >= 0, "create: bad oversampling");
  MMSB_CHECK_ARG(cfg->ncom + cfg->oversample <= std::min(std::min(cfg->na, cfg->nb), cfg->nc),
                 "create: k + oversample must not exceed the partition sizes");
  MMSB_CHECK_ARG(cfg->minibatch >= 1 && cfg->power_iters >= 0 && cfg->ns_iters >= 0, "create: bad STGD settings");
  MMSB_CHECK_ARG(cfg->ns_degree == 3 || cfg->ns_degree == 5, "create: ns_degree must be 3 or 5 (got %d)", cfg->ns_degree);
  MMSB_CHECK_ARG(cfg->y_mode >= MMSB_Y_AUTO && cfg->y_mode <= MMSB_Y_RECOMPUTE, "create: unknown y_mode %d", cfg->y_mode);
  mmsb_ctx* c = new (std::nothrow) mmsb_ctx();
  if (!c) return set_error(MMSB_ERR_NOMEM, "create: out of host memory");
  c->cfg = *cfg;
  set_sizes(c);
  hipError_t e = hipSetDevice(cfg->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return set_error(MMSB_ERR_HIP, "create: %s", hipGetErrorString(e));
  }
  *out = c;
  g_live_contexts.fetch_add(1);
  return MMSB_OK;
}

extern "C" int mmsb_set_learn_rate(mmsb_ctx* c, double eta) {
  clear_error();
  MMSB_CHECK_ARG(c, "set_learn_rate: null context");
  MMSB_CHECK_ARG(std::isfinite(eta) && eta > 0.0, "set_learn_rate: learning rate %g must be positive", eta);
  HCHK(hipSetDevice(c->cfg.device));
  TRY(drop_epoch_graphs(c));
  c->cfg.learn_rate = eta;
  return MMSB_OK;
}

extern "C" void mmsb_destroy(mmsb_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->cfg.device);
  if (c->s) (void)hipStreamSynchronize(c->s);
  for (auto& plane : c->epoch_graph)
    for (auto& row : plane)
      for (hipGraphExec_t& g : row)
        if (g) (void)hipGraphExecDestroy(g);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->s) (void)hipStreamDestroy(c->s);
  delete c;
  if (g_live_contexts.fetch_sub(1) == 1) release_process_resources();
}

extern "C" int mmsb_set_block(mmsb_ctx* c, int32_t which, const mmsb_csr* b) {
  clear_error();
  MMSB_CHECK_ARG(c && b && which >= 0 && which < 5, "set_block: bad arguments");
  const int32_t rows_of[5] = {c->cfg.nx, c->cfg.nx, c->cfg.nx, c->cfg.nb, c->cfg.nc};
  const int32_t cols_of[5] = {c->cfg.na, c->cfg.nb, c->cfg.nc, c->cfg.na, c->cfg.na};
  MMSB_CHECK_ARG(b->n_rows == rows_of[which] && b->n_cols == cols_of[which], "set_block %d: shape %d x %d, expected %d x %d",
                 which, b->n_rows, b->n_cols, rows_of[which], cols_of[which]);
  MMSB_CHECK_ARG(b->nnz >= 0 && b->nnz < INT32_MAX && b->rowptr && (b->nnz == 0 || b->colidx), "set_block: bad CSR");
  MMSB_CHECK_ARG(b->rowptr[0] == 0 && b->rowptr[b->n_rows] == b->nnz, "set_block: rowptr does not match nnz");
  {  // every row pointer monotone, every column index in range (threads report the first bad index)
    const int nt = default_threads();
    std::vector<int64_t> bad_row(nt, -1), bad_nz(nt, -1);
    parallel_chunks(b->n_rows, nt, [&](int t, int64_t lo, int64_t hi) {
      for (int64_t i = lo; i < hi && bad_row[t] < 0; ++i)
        if (b->rowptr[i + 1] < b->rowptr[i]) bad_row[t] = i;
    });
    for (int64_t r : bad_row) MMSB_CHECK_ARG(r < 0, "set_block: rowptr not monotone at row %lld", (long long)r);
    parallel_chunks(b->nnz, nt, [&](int t, int64_t lo, int64_t hi) {
      for (int64_t e = lo; e < hi && bad_nz[t] < 0; ++e)
        if (b->colidx[e] < 0 || b->colidx[e] >= b->n_cols) bad_nz[t] = e;
    });
    for (int64_t e : bad_nz)
      MMSB_CHECK_ARG(e < 0, "set_block: column index %d out of range", e < 0 ? 0 : b->colidx[e]);
  }
  HCHK(hipSetDevice(c->cfg.device));
  if (c->have[which]) return set_error(MMSB_ERR_STATE, "set_block: block %d already set", which);
  TRY(upload_csr(c, c->blk[which], b->n_rows, b->n_cols, b->nnz, b->rowptr, b->colidx, b->vals));
  TRY(transpose_csr(c, c->blk[which], c->blkT[which]));
  c->have[which] = true;
  return MMSB_OK;
}

extern "C" int mmsb_read_inputs(mmsb_ctx* c) {
  clear_error();
  MMSB_CHECK_ARG(c, "read_inputs: null context");
  const stpm_config& cf = c->cfg;
  const char* stems[5] = {"Gx_a", "Gx_b", "Gx_c", "Gb_a", "Gc_a"};
  const int32_t rows_of[5] = {cf.nx, cf.nx, cf.nx, cf.nb, cf.nc};
  const int32_t cols_of[5] = {cf.na, cf.nb, cf.nc, cf.na, cf.na};
  const int nfiles = cf.synthetic ? 5 : 3;
  for (int b = 0; b < nfiles; ++b) {
    const std::string path = std::string(cf.folder_name) + cf.file_prefix + stems[b] + ".txt";
    c->log("reading %s\n", path.c_str());
    mmsb_csr m{};
    const int32_t first_is_row = b < 3 ? 1 : cf.ba_first_is_row;
    const int st = mmsb_read_coo(path.c_str(), rows_of[b], cols_of[b], first_is_row, cf.binary ? 0 : 1,
                                 cf.binary ? MMSB_DUP_ONE : MMSB_DUP_LAST, cf.feof_repeat, &m);
    if (st != MMSB_OK) return st;
    const int st2 = mmsb_set_block(c, b, &m);
    mmsb_csr_free(&m);
    if (st2 != MMSB_OK) return st2;
  }
  return MMSB_OK;
}

extern "C" int mmsb_whiten(mmsb_ctx* c) {
  clear_error();
  MMSB_CHECK_ARG(c, "whiten: null context");
  for (int b = 0; b < 3; ++b)
    if (!c->have[b]) return set_error(MMSB_ERR_STATE, "whiten: block %d (G_X%c) not set", b, "ABC"[b]);
  HCHK(hipSetDevice(c->cfg.device));
  const double t0 = now_s();
  TRY(whiten_impl(c));
  HCHK(hipStreamSynchronize(c->s));
  c->times[0] = now_s() - t0;
  return MMSB_OK;
}

extern "C" int mmsb_set_phi(mmsb_ctx* c, const double* phi) {
  clear_error();
  MMSB_CHECK_ARG(c && phi, "set_phi: null argument");
  HCHK(hipSetDevice(c->cfg.device));
  TRY(alloc_stgd(c));
  std::vector<float> buf((size_t)c->kp * c->kp, 0.f);
  for (int32_t m = 0; m < c->k; ++m)
    for (int32_t i = 0; i < c->k; ++i) buf[(size_t)m * c->kp + i] = (float)phi[(size_t)m * c->k + i];
  HCHK(hipMemcpyAsync(c->phi[c->cur], buf.data(), sizeof(float) * buf.size(), hipMemcpyHostToDevice, c->s));
  HCHK(hipStreamSynchronize(c->s));
  c->have_phi = true;
  c->eig_done = c->post_done = false;
  return MMSB_OK;
}

extern "C" int mmsb_set_yvectors(mmsb_ctx* c, int64_t n_x, int32_t k, const double* ya, const double* yb, const double* yc,
                                 const double* mua, const double* mub, const double* muc) {
  clear_error();
  MMSB_CHECK_ARG(c && ya && yb && yc && mua && mub && muc, "set_yvectors: null argument");
  MMSB_CHECK_ARG(n_x == c->n[0] && k == c->k, "set_yvectors: shape %lld x %d, context has %lld x %d", (long long)n_x, k,
                 (long long)c->n[0], c->k);
  HCHK(hipSetDevice(c->cfg.device));
  if (c->y_window) TRY(drop_epoch_graphs(c));
  c->y_window = false;
  TRY(alloc_y(c));
  const int32_t kp = c->kp;
  std::vector<float> a((size_t)c->n_pad * kp, 0.f), b((size_t)c->n_pad * kp, 0.f), ct((size_t)kp * c->n_pad, 0.f);
  for (int64_t x = 0; x < n_x; ++x)
    for (int32_t i = 0; i < k; ++i) {
      a[(size_t)x * kp + i] = (float)ya[x * k + i];
      b[(size_t)x * kp + i] = (float)yb[x * k + i];
      ct[(size_t)i * c->n_pad + x] = (float)yc[x * k + i];
    }
  TRY(h2d(c->s, c->YA, a.data(), sizeof(float) * a.size()));
  TRY(h2d(c->s, c->YB, b.data(), sizeof(float) * b.size()));
  TRY(h2d(c->s, c->YCt, ct.data(), sizeof(float) * ct.size()));
  const double* mus[3] = {mua, mub, muc};
  for (int i = 0; i < 3; ++i) HCHK(hipMemcpyAsync(c->muw[i], mus[i], sizeof(double) * k, hipMemcpyHostToDevice, c->s));
  HCHK(hipStreamSynchronize(c->s));
  c->have_y = true;
  c->y16_valid = false;
  return MMSB_OK;
}

extern "C" int mmsb_stgd_timed(mmsb_ctx* c, int64_t steps, int32_t precision, double* phase_ms) {
  clear_error();
  MMSB_CHECK_ARG(c && phase_ms, "stgd_timed: null argument");
  MMSB_CHECK_ARG(steps >= 0, "stgd_timed: negative step count");
  if (!c->have_y && !(c->y_window && c->whitened))
    return set_error(MMSB_ERR_STATE, "stgd_timed: neighbour vectors not computed (call mmsb_whiten)");
  HCHK(hipSetDevice(c->cfg.device));
  TRY(alloc_stgd(c));
  if (!c->have_phi) TRY(init_phi(c));
  c->eig_done = c->post_done = false;
  StgdTiming tm{};
  int st = MMSB_OK;
  for (int i = 0; i < 4 && st == MMSB_OK; ++i)
    if (hipEventCreate(&tm.ev[i]) != hipSuccess) st = set_error(MMSB_ERR_HIP, "stgd_timed: hipEventCreate failed");
  // the cost of an event pair with nothing between them, subtracted from every phase interval
  double empty_ms = 0.0;
  const int kEmpty = 20;
  for (int i = 0; i < kEmpty && st == MMSB_OK; ++i) {
    float ms = 0.f;
    if (hipEventRecord(tm.ev[0], c->s) != hipSuccess || hipEventRecord(tm.ev[1], c->s) != hipSuccess ||
        hipEventSynchronize(tm.ev[1]) != hipSuccess || hipEventElapsedTime(&ms, tm.ev[0], tm.ev[1]) != hipSuccess)
      st = set_error(MMSB_ERR_HIP, "stgd_timed: empty event pair failed");
    empty_ms += ms;
  }
  empty_ms /= kEmpty;
  if (st == MMSB_OK) st = run_stgd_steps(c, steps, precision, &tm);
  for (int i = 0; i < 4; ++i)
    if (tm.ev[i]) (void)hipEventDestroy(tm.ev[i]);
  TRY(st);
  for (int i = 0; i < 3; ++i) phase_ms[i] = std::max(0.0, (double)tm.ms[i] - empty_ms * (double)steps);
  return MMSB_OK;
}

extern "C" int mmsb_stgd(mmsb_ctx* c, int64_t steps, int32_t precision) {
  clear_error();
  MMSB_CHECK_ARG(c, "stgd: null context");
  if (!c->have_y && !(c->y_window && c->whitened))
    return set_error(MMSB_ERR_STATE, "stgd: neighbour vectors not computed (call mmsb_whiten)");
  HCHK(hipSetDevice(c->cfg.device));
  TRY(alloc_stgd(c));
  if (!c->have_phi) TRY(init_phi(c));
  c->eig_done = c->post_done = false;
  const double t0 = now_s();
  if (steps >= 0) {
    TRY(run_stgd_steps(c, steps, precision));
  } else {
    // The configured epochs, each ended early by the reference's test after any step (stpm.c:747-765:
    // ||dPhi||^2 < THRESH after more than SGD_ITER_MIN steps, or more than SGD_ITER_MAX steps) or by a
    // stop request at an epoch boundary (stpm.c:754).
    const stpm_config& cf = c->cfg;
    const int64_t spe = (c->n[0] + c->bl - 1) / c->bl;
    const bool rule = cf.sgd_iter_min > 0 || cf.sgd_iter_max > 0;
    for (int32_t e = 0; e < cf.epochs; ++e) {
      const bool low = precision == MMSB_PREC_BF16 && e < cf.epochs - cf.polish_epochs;
      bool stopped = false;
      TRY(run_stgd_steps(c, spe, low ? MMSB_PREC_BF16 : MMSB_PREC_FP32, nullptr, rule, &stopped));
      if (stopped) {
        c->log("STGD loop ended after step %lld\n", (long long)c->steps_done);
        break;
      }
      if (e + 1 < cf.epochs) {
        HCHK(hipStreamSynchronize(c->s));
        if (c->stop_req.load()) {
          c->log("Exiting STPM loop now\n");
          break;
        }
      }
    }
  }
  HCHK(hipStreamSynchronize(c->s));
  c->stop_req.store(0);
  c->times[1] += now_s() - t0;
  return MMSB_OK;
}

extern "C" int mmsb_request_stop(mmsb_ctx* c) {
  clear_error();
  MMSB_CHECK_ARG(c, "request_stop: null context");
  c->stop_req.store(1);
  return MMSB_OK;
}

extern "C" int mmsb_grad(mmsb_ctx* c, int64_t row0, int32_t nrows, const double* phi, int32_t scaling, double theta,
                         int32_t recompute, double* grad, double* loss) {
  clear_error();
  MMSB_CHECK_ARG(c && phi && grad && loss, "grad: null argument");
  MMSB_CHECK_ARG(nrows >= 1 && row0 >= 0 && row0 + nrows <= c->n[0], "grad: rows [%lld, %lld) outside X", (long long)row0,
                 (long long)(row0 + nrows));
  if (!recompute && !c->have_y && c->y_window) recompute = 1;  // nothing resident to read
  if (recompute && !c->whitened) return set_error(MMSB_ERR_STATE, "grad: recomputing needs the whitening factors");
  if (!recompute && !c->have_y) return set_error(MMSB_ERR_STATE, "grad: neighbour vectors not available");
  HCHK(hipSetDevice(c->cfg.device));
  hipStream_t s = c->s;
  const int32_t k = c->k, kp = c->kp;
  float *ya = nullptr, *yb = nullptr, *yc = nullptr;
  double *dphi = nullptr, *dgrad = nullptr, *dloss = nullptr;
  unsigned char* work = nullptr;
  const size_t wb = grad_workspace_bytes(k, nrows);
  // scratch of this call (freed before returning)
  auto cleanup = [&]() {
    for (void* p : {(void*)ya, (void*)yb, (void*)yc, (void*)dphi, (void*)dgrad, (void*)dloss, (void*)work})
      if (p) (void)hipFree(p);
  };
  hipError_t e = hipSuccess;
  const size_t ybytes = sizeof(float) * (size_t)nrows * kp;
  for (float** p : {&ya, &yb, &yc})
    if (e == hipSuccess) e = hipMalloc(p, ybytes);
  if (e == hipSuccess) e = hipMalloc(&dphi, sizeof(double) * k * k);
  if (e == hipSuccess) e = hipMalloc(&dgrad, sizeof(double) * k * k);
  if (e == hipSuccess) e = hipMalloc(&dloss, sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&work, wb);
  if (e != hipSuccess) {
    cleanup();
    return set_error(MMSB_ERR_NOMEM, "grad: %s", hipGetErrorString(e));
  }
  int st = MMSB_OK;
  if (recompute) {  // y rows = (G_X* rows) x (W, Z_B^T W, Z_C^T W)
    const DevCsr* g[3] = {&c->blk[0], &c->blk[1], &c->blk[2]};
    const float* rhs[3] = {c->W, c->MB, c->MC};
    float* out[3] = {ya, yb, yc};
    for (int i = 0; i < 3 && st == MMSB_OK; ++i) st = spmm_rows(c, s, *g[i], row0, nrows, rhs[i], kp, k, out[i], kp);
  } else {
    e = hipMemcpy2DAsync(ya, sizeof(float) * kp, c->YA + row0 * kp, sizeof(float) * kp, sizeof(float) * kp, nrows,
                         hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpy2DAsync(yb, sizeof(float) * kp, c->YB + row0 * kp, sizeof(float) * kp, sizeof(float) * kp, nrows,
                           hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) st = set_error(MMSB_ERR_HIP, "grad: %s", hipGetErrorString(e));
    if (st == MMSB_OK) st = transpose_f32(s, k, nrows, c->YCt + row0, c->n_pad, yc, kp);  // [k x n] -> [n x k]
  }
  double co[3];
  if (st == MMSB_OK) st = mmsb_tensor_coeffs(c->cfg.alpha_0, scaling, co);
  if (st == MMSB_OK) {
    e = hipMemcpyAsync(dphi, phi, sizeof(double) * k * k, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) st = set_error(MMSB_ERR_HIP, "grad: %s", hipGetErrorString(e));
  }
  if (st == MMSB_OK)
    st = stgd_grad<float>(s, k, nrows, ya, yb, yc, kp, c->muw[0], c->muw[1], c->muw[2], dphi, co[0], co[1], co[2], theta, dgrad,
                          dloss, work);
  if (st == MMSB_OK) {
    e = hipMemcpyAsync(grad, dgrad, sizeof(double) * k * k, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(loss, dloss, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) st = set_error(MMSB_ERR_HIP, "grad: %s", hipGetErrorString(e));
  } else {
    (void)hipStreamSynchronize(s);
  }
  cleanup();
  return st;
}

extern "C" int mmsb_power_method(mmsb_ctx* c, int32_t iters, int32_t orthogonalize, double* lambda) {
  clear_error();
  MMSB_CHECK_ARG(c && lambda && iters >= 0, "power_method: bad arguments");
  if (!c->have_y && !(c->y_window && c->whitened))
    return set_error(MMSB_ERR_STATE, "power_method: neighbour vectors not available");
  HCHK(hipSetDevice(c->cfg.device));
  TRY(alloc_stgd(c));
  if (!c->have_phi) TRY(init_phi(c));
  const int32_t k = c->k, kp = c->kp;
  const int32_t chunk = (int32_t)std::min<int64_t>(c->n[0], 65536);
  double co[3];
  TRY(mmsb_tensor_coeffs(c->cfg.alpha_0, MMSB_M2_PAPER, co));
  // Phi (FP32, kp-padded) <-> compact FP64 k x k
  std::vector<float> hf((size_t)kp * kp);
  std::vector<double> hd((size_t)k * k);
  HCHK(hipMemcpy(hf.data(), c->phi[c->cur], sizeof(float) * hf.size(), hipMemcpyDeviceToHost));
  for (int32_t m = 0; m < k; ++m)
    for (int32_t i = 0; i < k; ++i) hd[(size_t)m * k + i] = hf[(size_t)m * kp + i];
  double *dphi = nullptr, *dlam = nullptr;
  float* yr[3] = {nullptr, nullptr, nullptr};  // a chunk of Y_C rows (resident form) or of all three (recompute form)
  void* work = nullptr;
  auto cleanup = [&]() {
    for (void* p : {(void*)dphi, (void*)dlam, (void*)yr[0], (void*)yr[1], (void*)yr[2], work})
      if (p) (void)hipFree(p);
  };
  hipError_t e = hipMalloc(&dphi, sizeof(double) * k * k);
  if (e == hipSuccess) e = hipMalloc(&dlam, sizeof(double) * k);
  for (int i = 0; i < (c->have_y ? 1 : 3); ++i)
    if (e == hipSuccess) e = hipMalloc(&yr[i], sizeof(float) * (size_t)chunk * kp);
  if (e == hipSuccess) e = hipMalloc(&work, power_workspace_bytes(k, chunk));
  if (e == hipSuccess) e = hipMemcpyAsync(dphi, hd.data(), sizeof(double) * k * k, hipMemcpyHostToDevice, c->s);
  if (e != hipSuccess) {
    cleanup();
    return set_error(MMSB_ERR_NOMEM, "power_method: %s", hipGetErrorString(e));
  }
  PowerRows rows;
  if (c->have_y) {  // resident: Y_A, Y_B rows in place, Y_C^T columns transposed into a chunk
    rows = [c, &yr](int64_t r0, int32_t nb, const float** ya, const float** yb, const float** yc) -> int {
      TRY(transpose_f32(c->s, c->k, nb, c->YCt + r0, c->n_pad, yr[0], c->kp));  // [k x nb] -> [nb x k]
      *ya = c->YA + r0 * c->kp;
      *yb = c->YB + r0 * c->kp;
      *yc = yr[0];
      return MMSB_OK;
    };
  } else {  // recompute form: the chunk's rows rebuilt from CSR as STGD's minibatches are
    rows = [c, &yr](int64_t r0, int32_t nb, const float** ya, const float** yb, const float** yc) -> int {
      YRowsJobs jobs{};
      const float* R[3] = {c->W, c->MB, c->MC};
      for (int i = 0; i < 3; ++i) {
        const DevCsr& g = c->blk[i];
        jobs.j[i] = YRowsJob{g.rowptr + r0, g.colidx, g.vals, R[i], yr[i], c->kp, 0, (int64_t)g.n_cols};
      }
      jobs.count = 3;
      TRY(spmm_yrows(c->s, jobs, nb, c->k, c->kp, false));
      *ya = yr[0];
      *yb = yr[1];
      *yc = yr[2];
      return MMSB_OK;
    };
  }
  int st = MMSB_OK;
  for (int32_t it = 0; it < iters && st == MMSB_OK; ++it)
    st = power_iteration(c->s, k, c->n[0], rows, kp, c->muw[0], c->muw[1], c->muw[2], co[0], co[1], co[2], dphi, dlam,
                         orthogonalize, chunk, work);
  if (st == MMSB_OK) {
    e = hipMemcpyAsync(hd.data(), dphi, sizeof(double) * k * k, hipMemcpyDeviceToHost, c->s);
    if (e == hipSuccess) e = hipMemcpyAsync(lambda, dlam, sizeof(double) * k, hipMemcpyDeviceToHost, c->s);
    if (e == hipSuccess) e = hipStreamSynchronize(c->s);
    if (e != hipSuccess) st = set_error(MMSB_ERR_HIP, "power_method: %s", hipGetErrorString(e));
  } else {
    (void)hipStreamSynchronize(c->s);
  }
  cleanup();
  if (st != MMSB_OK) return st;
  if (iters == 0) std::fill(lambda, lambda + k, 0.0);
  return mmsb_set_phi(c, hd.data());
}

extern "C" int mmsb_eigen(mmsb_ctx* c) {
  clear_error();
  MMSB_CHECK_ARG(c, "eigen: null context");
  if (!c->have_phi) return set_error(MMSB_ERR_STATE, "eigen: no Phi yet");
  HCHK(hipSetDevice(c->cfg.device));
  TRY(eigen_impl(c));
  HCHK(hipStreamSynchronize(c->s));
  return MMSB_OK;
}

extern "C" int mmsb_post(mmsb_ctx* c) {
  clear_error();
  MMSB_CHECK_ARG(c, "post: null context");
  if (!c->have_phi) return set_error(MMSB_ERR_STATE, "post: no Phi yet");
  HCHK(hipSetDevice(c->cfg.device));
  const double t0 = now_s();
  TRY(post_impl(c));
  HCHK(hipStreamSynchronize(c->s));
  c->times[2] = now_s() - t0;
  c->times[3] = c->times[0] + c->times[1] + c->times[2];
  return MMSB_OK;
}

namespace mmsb {
// Y_A / Y_B / Y_C rows of a context that recomputes its minibatches: rebuilt from CSR by the STGD
// fill kernel (FP32) in chunks of rows, [nrows x k] row-major doubles.
static bool y_recomputed(const mmsb_ctx* c, int32_t what) {
  return what >= MMSB_Q_YA && what <= MMSB_Q_YC && !c->have_y && c->y_window && c->whitened;
}

static int yrows_host(mmsb_ctx* c, int32_t what, int64_t row0, int64_t nrows, double* host) {
  const int i = what - MMSB_Q_YA;
  const int32_t k = c->k, kp = c->kp;
  const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(nrows, 65536));
  float* d = nullptr;
  HCHK(hipMalloc(&d, sizeof(float) * (size_t)chunk * kp));
  std::vector<float> h((size_t)chunk * kp);
  const float* R[3] = {c->W, c->MB, c->MC};
  const DevCsr& g = c->blk[i];
  int st = MMSB_OK;
  for (int64_t r = 0; r < nrows && st == MMSB_OK; r += chunk) {
    const int32_t m = (int32_t)std::min<int64_t>(chunk, nrows - r);
    YRowsJobs jobs{};
    jobs.j[0] = YRowsJob{g.rowptr + row0 + r, g.colidx, g.vals, R[i], d, kp, 0, (int64_t)g.n_cols};
    jobs.count = 1;
    st = spmm_yrows(c->s, jobs, m, k, kp, false);
    if (st == MMSB_OK) st = d2h(c->s, h.data(), d, sizeof(float) * (size_t)m * kp);
    if (st == MMSB_OK)
      for (int64_t x = 0; x < m; ++x)
        for (int32_t j = 0; j < k; ++j) host[(r + x) * k + j] = h[(size_t)x * kp + j];
  }
  (void)hipFree(d);
  return st;
}
}  // namespace mmsb

extern "C" int64_t mmsb_get_size(mmsb_ctx* c, int32_t what) {
  if (!c) return 0;
  if (what == MMSB_Q_TIMES) return 4;
  if (what == MMSB_Q_SWEEPS) return c->whitened ? 2 : 0;
  if (y_recomputed(c, what)) return c->n[0] * c->k;
  const View v = view_of(c, what);
  return v.p ? v.rows * v.cols : 0;
}

extern "C" int mmsb_get(mmsb_ctx* c, int32_t what, double* host, int64_t count) {
  clear_error();
  MMSB_CHECK_ARG(c && host, "get: null argument");
  if (what == MMSB_Q_TIMES) {
    MMSB_CHECK_ARG(count >= 4, "get: need 4 doubles for the timings");
    std::memcpy(host, c->times, sizeof(c->times));
    return MMSB_OK;
  }
  if (what == MMSB_Q_SWEEPS) {
    MMSB_CHECK_ARG(c->whitened, "get: no whitening has run");
    MMSB_CHECK_ARG(count >= 2, "get: need 2 doubles for the sweep counts");
    int32_t st[8];
    HCHK(hipSetDevice(c->cfg.device));
    HCHK(hipMemcpyAsync(st, c->dstat, sizeof(st), hipMemcpyDeviceToHost, c->s));
    HCHK(hipStreamSynchronize(c->s));
    host[0] = st[6];
    host[1] = st[7];
    return MMSB_OK;
  }
  if (y_recomputed(c, what)) {
    MMSB_CHECK_ARG(count >= c->n[0] * c->k, "get: buffer too small");
    HCHK(hipSetDevice(c->cfg.device));
    return yrows_host(c, what, 0, c->n[0], host);
  }
  const View v = view_of(c, what);
  if (!v.p) return set_error(MMSB_ERR_STATE, "get: quantity %d not computed", what);
  MMSB_CHECK_ARG(count >= v.rows * v.cols, "get: buffer of %lld doubles, need %lld", (long long)count, (long long)(v.rows * v.cols));
  HCHK(hipSetDevice(c->cfg.device));
  return copy_out(c, v, host);
}

extern "C" int mmsb_get_rows(mmsb_ctx* c, int32_t what, int64_t row0, int64_t nrows, double* host, int64_t count) {
  clear_error();
  MMSB_CHECK_ARG(c && host, "get_rows: null argument");
  if (y_recomputed(c, what)) {
    MMSB_CHECK_ARG(row0 >= 0 && nrows >= 0 && row0 + nrows <= c->n[0], "get_rows: rows [%lld, %lld) outside [0, %lld)",
                   (long long)row0, (long long)(row0 + nrows), (long long)c->n[0]);
    MMSB_CHECK_ARG(count >= nrows * c->k, "get_rows: buffer too small");
    HCHK(hipSetDevice(c->cfg.device));
    return yrows_host(c, what, row0, nrows, host);
  }
  View v = view_of(c, what);
  if (!v.p) return set_error(MMSB_ERR_STATE, "get_rows: quantity %d not computed", what);
  MMSB_CHECK_ARG(row0 >= 0 && nrows >= 0 && row0 + nrows <= v.rows, "get_rows: rows [%lld, %lld) outside [0, %lld)",
                 (long long)row0, (long long)(row0 + nrows), (long long)v.rows);
  MMSB_CHECK_ARG(count >= nrows * v.cols, "get_rows: buffer too small");
  const size_t esz = v.dtype == 1 ? 8 : 4;
  if (v.trans)
    v.p = static_cast<const unsigned char*>(v.p) + row0 * esz;
  else
    v.p = static_cast<const unsigned char*>(v.p) + row0 * v.ld * esz;
  v.rows = nrows;
  HCHK(hipSetDevice(c->cfg.device));
  return copy_out(c, v, host);
}

extern "C" int mmsb_get_f32(mmsb_ctx* c, int32_t what, float* host, int64_t count) {
  clear_error();
  MMSB_CHECK_ARG(c && host, "get_f32: null argument");
  const View v = view_of(c, what);
  if (!v.p) return set_error(MMSB_ERR_STATE, "get_f32: quantity %d not computed", what);
  MMSB_CHECK_ARG(v.dtype == 0 && !v.trans, "get_f32: quantity %d is not a row-major FP32 matrix", what);
  MMSB_CHECK_ARG(count >= v.rows * v.cols, "get_f32: buffer too small");
  HCHK(hipSetDevice(c->cfg.device));
  return copy_raw(c, v, host);
}

extern "C" int mmsb_write_triplets(mmsb_ctx* c, const char* pi_path, const char* white_path) {
  clear_error();
  MMSB_CHECK_ARG(c, "write_triplets: null context");
  HCHK(hipSetDevice(c->cfg.device));
  return write_triplets_impl(c, pi_path, white_path);
}

extern "C" int mmsb_sync(mmsb_ctx* c) {
  clear_error();
  MMSB_CHECK_ARG(c, "sync: null context");
  HCHK(hipStreamSynchronize(c->s));
  return MMSB_OK;
}

extern "C" int mmsb_write_outputs(mmsb_ctx* c) {
  clear_error();
  MMSB_CHECK_ARG(c, "write_outputs: null context");
  const stpm_config& cf = c->cfg;
  const char* pis[4] = {"Pi_X", "Pi_A", "Pi_B", "Pi_C"};
  for (int p = 0; p < 4; ++p) {
    const View v = view_of(c, MMSB_Q_PI_X + p);
    if (!v.p) continue;
    std::vector<double> h((size_t)v.rows * v.cols);
    TRY(copy_out(c, v, h.data()));
    TRY(mmsb_write_values(out_name(cf, pis[p]).c_str(), h.data(), (int64_t)h.size()));  // post_proc.c:110-114
  }
  if (c->have_phi) {  // evecs: k x k column-major (stpm.c:935-937)
    std::vector<double> h((size_t)c->k * c->k), cm((size_t)c->k * c->k);
    TRY(copy_out(c, view_of(c, MMSB_Q_PHI), h.data()));
    for (int32_t i = 0; i < c->k; ++i)
      for (int32_t m = 0; m < c->k; ++m) cm[(size_t)i * c->k + m] = h[(size_t)m * c->k + i];
    TRY(mmsb_write_values(out_name(cf, "evecs").c_str(), cm.data(), (int64_t)cm.size()));
  }
  if (c->eig_done) {  // evals: alpha_hat (stpm.c:926-929)
    std::vector<double> h(c->k);
    TRY(copy_out(c, view_of(c, MMSB_Q_ALPHA_HAT), h.data()));
    TRY(mmsb_write_values(out_name(cf, "evals").c_str(), h.data(), c->k));
  }
  if (c->steps_done) {  // frob_norm (stpm.c:948-951)
    std::vector<double> h((size_t)c->steps_done);
    TRY(copy_out(c, view_of(c, MMSB_Q_FROB), h.data()));
    TRY(mmsb_write_values(out_name(cf, "frob_norm").c_str(), h.data(), c->steps_done));
  }
  // the CPU program's normalised-Pi and W triplet files (CommunityDetection.cpp:141-142)
  TRY(write_triplets_impl(c, c->post_done ? out_name(cf, "pi").c_str() : nullptr,
                          c->whitened ? out_name(cf, "white").c_str() : nullptr));
  const std::string path = out_name(cf, "sim_info");  // stpm.c:957-975
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) return set_error(MMSB_ERR_IO, "writing %s failed", path.c_str());
  std::fprintf(f, "Total number of nodes          = %d\n", cf.nx + cf.na + cf.nb + cf.nc);
  std::fprintf(f, "Number of nodes in partition X = %d\n", cf.nx);
  std::fprintf(f, "Number of nodes in partition A = %d\n", cf.na);
  std::fprintf(f, "Number of nodes in partition B = %d\n", cf.nb);
  std::fprintf(f, "Number of nodes in partition C = %d\n", cf.nc);
  std::fprintf(f, "Number of nodes communities    = %d\n", cf.ncom);
  std::fprintf(f, "Alpha_0 (Dirichlet parameter)  = %5.25e\n", cf.alpha_0);
  std::fprintf(f, "Learning rate = %5.25e\n", cf.learn_rate);
  std::fprintf(f, "Threshold for STPM convergence    = %5.25e\n", cf.thresh);
  std::fprintf(f, "Pseudoinverse numerical tolerance = %5.25e\n", cf.pinv_tol);
  std::fprintf(f, "Total number of Frob Norm values written to file = %d\n", (int)c->steps_done);
  std::fprintf(f, "Number of iterations between convergence test    = %d\n", 1);
  std::fprintf(f, "Maximum number of stochastic iterations allowed  = %d\n", (int)cf.sgd_iter_max);
  std::fprintf(f, "Exec Time svd1  = %5.25e (Seconds)\n", c->t_svd[0]);
  std::fprintf(f, "Exec Time svd2  = %5.25e (Seconds)\n", c->t_svd[1]);
  std::fprintf(f, "Exec Time Pre Proc  = %5.25e (Seconds)\n", c->times[0]);
  std::fprintf(f, "Exec Time STPM      = %5.25e (Seconds)\n", c->times[1]);
  std::fprintf(f, "Exec Time Post Proc = %5.25e (Seconds)\n", c->times[2]);
  std::fprintf(f, "Total Exec Time     = %5.25e (Seconds)\n", c->times[0] + c->times[1] + c->times[2]);
  if (std::fclose(f) != 0) return set_error(MMSB_ERR_IO, "closing %s failed", path.c_str());
  return MMSB_OK;
}

// ---- program-level drop-in: one process-wide context, as the reference's globals ------------------
static stpm_config g_cfg;
static bool g_cfg_set = false;
static mmsb_ctx* g_ctx = nullptr;
static int g_task_status = 0;

extern "C" int stpm_set_config(const stpm_config* cfg) {
  clear_error();
  MMSB_CHECK_ARG(cfg, "stpm_set_config: null");
  g_cfg = *cfg;
  g_cfg_set = true;
  if (g_ctx) {
    mmsb_destroy(g_ctx);
    g_ctx = nullptr;
  }
  return MMSB_OK;
}

extern "C" int pre_proc(void) {
  clear_error();
  if (!g_cfg_set) stpm_config_default(&g_cfg), g_cfg_set = true;
  if (g_ctx) mmsb_destroy(g_ctx), g_ctx = nullptr;
  TRY(mmsb_create(&g_cfg, &g_ctx));
  TRY(mmsb_read_inputs(g_ctx));
  return mmsb_whiten(g_ctx);
}

extern "C" int post_proc(void) {
  clear_error();
  if (!g_ctx) return set_error(MMSB_ERR_STATE, "post_proc before pre_proc");
  TRY(mmsb_post(g_ctx));
  return mmsb_write_outputs(g_ctx);
}

extern "C" void stpm_force_exit(void) {
  mmsb_ctx* c = g_ctx;
  if (c) c->stop_req.store(1);
}

extern "C" void mmsb_release(void) {
  if (g_ctx) {
    mmsb_destroy(g_ctx);
    g_ctx = nullptr;
  }
  release_process_resources();
}

extern "C" void* GPU_Task(void*) {
  g_task_status = 0;
  if (!g_ctx) {
    g_task_status = set_error(MMSB_ERR_STATE, "GPU_Task before pre_proc");
    return &g_task_status;
  }
  int st = mmsb_stgd(g_ctx, -1, g_cfg.precision);
  if (st == MMSB_OK) st = mmsb_eigen(g_ctx);
  if (st == MMSB_OK) st = post_proc();
  g_task_status = st;
  if (st != MMSB_OK) std::fprintf(stderr, "GPU_Task: %s\n", mmsb_last_error());  // the message is thread-local
  return st == MMSB_OK ? nullptr : &g_task_status;
}
