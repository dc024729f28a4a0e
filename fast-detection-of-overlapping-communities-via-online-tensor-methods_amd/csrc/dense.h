// Internal declarations of the dense FP64 kernels (dense.hip, jacobi.hip).
#pragma once

#include <hip/hip_runtime_api.h>
#include <cstddef>
#include <cstdint>
#include <functional>

namespace mmsb {

int gaussian_fill(hipStream_t s, uint64_t seed, uint32_t index, int64_t n, int32_t r, float* out, int32_t ld);
int csr_rowsum(hipStream_t s, int32_t m, const int32_t* rowptr, const float* vals, double scale, double* out);

size_t tn_workspace_bytes(int64_t n, int32_t p, int32_t q);
// C[p x q] = alpha * A^T B + beta * C (FP64 accumulate; `work` holds tn_workspace_bytes).
template <typename TA, typename TB>
int gemm_tn(hipStream_t s, int64_t n, int32_t p, int32_t q, double alpha, const TA* A, int32_t lda, const TB* B,
            int32_t ldb, double beta, double* C, double* work);
// Y[n x q] = alpha * X C + beta * Y, C FP64 [p x q] (row-major, or column-major when c_trans).
// Y = alpha X C + beta Y in FP32 on MFMA (C FP64 rounded to FP32, FP32 sums): for well-conditioned
// factors only.
int gemm_nn_f32(hipStream_t s, int64_t n, int32_t p, int32_t q, double alpha, const float* X, int32_t ldx,
                const double* C, int32_t ldc, double beta, float* Y, int32_t ldy, bool c_trans = false);
// C (FP64) = alpha A^T B + beta C with FP32 MFMA sums over row slices reduced in FP64; work as for
// gemm_tn (tn_workspace_bytes).
int gemm_tn_f32(hipStream_t s, int64_t n, int32_t p, int32_t q, double alpha, const float* A, int32_t lda, const float* B,
                int32_t ldb, double beta, double* C, double* work);
template <typename TX, typename TY>
int gemm_nn(hipStream_t s, int64_t n, int32_t p, int32_t q, double alpha, const TX* X, int32_t ldx, const double* C,
            int32_t ldc, double beta, TY* Y, int32_t ldy, bool c_trans = false);

// Upper Cholesky factor R of G + shift*trace(G)*I (G overwritten) and R^{-1}; status[0] != 0 on failure.
int chol_inv(hipStream_t s, int32_t r, double* G, double shift, double* R, double* Rinv, int32_t* status);

template <typename T>
int col_sign(hipStream_t s, int64_t n, int32_t c, const T* M, int32_t ld, unsigned long long* best, double* sign);
template <typename T>
int scale_cols(hipStream_t s, int64_t n, int32_t c, T* M, int32_t ld, const double* sc);

// One-sided Jacobi SVD of a square r x r FP64 matrix given column-major in A (overwritten):
// A = U diag(sigma) V^T with sigma sorted descending; U, V column-major r x r. V may be null when only
// U and sigma are wanted (e.g. the eigenvectors of a symmetric positive definite A): the rotations are
// then not accumulated. work: jacobi_workspace_bytes(r). sweeps_out[0] = sweeps used (device int).
size_t jacobi_workspace_bytes(int32_t r);
int jacobi_svd(hipStream_t s, int32_t r, double* A, double* U, double* sigma, double* V, void* work, int32_t* sweeps_out);

// G^T in CSR (rpT [nc + 1], ciT [nnz], vT [nnz] if v) of the device CSR G [nr x nc]; returns after
// the result is complete (csr_transpose.hip).
// order[0..nr) = the rows of a CSR matrix sorted by descending length (ties in ascending row order):
// a stable radix sort of (length, row) pairs; temporaries are allocated and freed inside, and the
// call returns once the order is complete (csr_transpose.hip).
int csr_row_order(hipStream_t s, int32_t nr, const int32_t* rp, int32_t* order);
int csr_transpose(hipStream_t s, int32_t nr, int32_t nc, int64_t nnz, const int32_t* rp, const int32_t* ci, const float* v,
                  int32_t* rpT, int32_t* ciT, float* vT);

int transpose_f32(hipStream_t s, int64_t R, int32_t C, const float* in, int64_t ldi, float* out, int64_t ldo);
int pinv_scale(hipStream_t s, int32_t k, const double* v, double shift, double tol, int mode, double* out, double* vout);
int sym_shift(hipStream_t s, int32_t r, double* S, double* shift_out);
int scale_rows(hipStream_t s, int32_t k, int32_t c, double* P, int32_t ld, const double* d, double alpha);
int diag_axpy(hipStream_t s, int64_t n, int32_t c, const double* mu, double alpha, const float* X, int32_t ldx, float* O,
              int32_t ldo);
int copy2d_f32(hipStream_t s, int32_t rows, int32_t cols, const float* in, int32_t ldi, float* out, int32_t ldo);
int spmm_f32(hipStream_t s, int32_t m, int32_t n, int64_t nnz, const int32_t* rowptr, const int32_t* colidx, const float* vals,
             float alpha, const float* B, int32_t ldb, int32_t r, float* Y, int32_t ldy, const int32_t* order = nullptr);
// The same product with the CSR block in host memory, streamed in row chunks of at most chunk_nnz
// nonzeros (spmm_stream.cpp); returns after the product has finished. check = false skips the scan of
// the row pointers and column indices, for blocks the caller has already validated.
int spmm_f32_host_csr(hipStream_t s, int32_t m, int32_t n, int64_t nnz, const int32_t* rowptr, const int32_t* colidx,
                      const float* vals, float alpha, const float* B, int32_t ldb, int32_t r, float* Y, int32_t ldy,
                      int64_t chunk_nnz, bool check = true);

// Whitened rows of one STGD minibatch: for each of the three jobs (y_A, y_B, y_C), rows [0, nb) of
// the CSR block whose rowptr is given (offset to the minibatch's first row) times R [n_cols x ldr]
// (r columns), written row-major ([nb x ldo]) or transposed ([r x ldo], trans != 0) as FP32 or BF16.
// Each row is summed in the order spmm_f32 uses and rounded like f32_to_bf16.
struct YRowsJob {
  const int32_t* rowptr;
  const int32_t* colidx;
  const float* vals;  // null: binary block
  const float* R;
  void* out;
  int64_t ldo;
  int32_t trans;
  int64_t rows_R;     // rows of R (its n_cols block dimension): sizes the gather for the strip choice
};
struct YRowsJobs {
  YRowsJob j[3];
  int32_t count;  // jobs used (1..3)
};
int spmm_yrows(hipStream_t s, const YRowsJobs& jobs, int32_t nb, int32_t r, int32_t ldr, bool bf16);

size_t grad_workspace_bytes(int32_t k, int32_t nb);
template <typename TY>
int stgd_grad(hipStream_t s, int32_t k, int32_t nb, const TY* ya, const TY* yb, const TY* yc, int32_t ldy,
              const double* mua, const double* mub, const double* muc, const double* phi, double c3, double c1, double c0,
              double theta, double* grad, double* loss, void* work);

size_t power_workspace_bytes(int32_t k, int32_t chunk);
// rows(r0, nb, &ya, &yb, &yc) enqueues whatever makes rows [r0, r0 + nb) of Y_A, Y_B, Y_C available
// row-major with leading dimension ldy, and points ya, yb, yc at them (valid until the next call).
using PowerRows = std::function<int(int64_t r0, int32_t nb, const float** ya, const float** yb, const float** yc)>;
int power_iteration(hipStream_t s, int32_t k, int64_t n, const PowerRows& rows, int32_t ldy, const double* mua,
                    const double* mub, const double* muc, double c3, double c1, double c0, double* phi, double* lam,
                    int orthogonalize, int32_t chunk, void* work);

}  // namespace mmsb
