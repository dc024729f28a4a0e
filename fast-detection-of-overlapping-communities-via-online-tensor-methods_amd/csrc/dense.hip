// Dense building blocks of the whitening stage, FP64 arithmetic on gfx950.
//
// The reference runs every dense product through culaBlockDgemm (gemm_pinv_nys.c:27-145) on
// n x n buffers and factors with culaDgeqrfp/culaDorgqr/culaDgesvd (gemm_pinv_nys.c:284-299,
// 385-394). Here the dense operands are tall n x r (r = k + oversampling) or small r x r, so the
// work is: tall-skinny A^T B (Gram-type, reduce over n), tall X * small C, a Cholesky factor with
// its triangular inverse (CholeskyQR), the Gaussian test matrices and column-sign canonicalisation.
// Tall operands are FP32 in HBM (the SpMM outputs) and are widened to FP64 on load; the
// accumulation and every small matrix are FP64 -- MI355X runs vector FP64 FMA at the same
// 78.6 TFLOP/s as its FP64 matrix cores, so these kernels use plain FP64 FMAs.
#include <hip/hip_cooperative_groups.h>

#include <type_traits>

#include "common.h"

#include "dense.h"
#include "philox.h"

namespace mmsb {

// ---- Gaussian test matrix (replaces gen_rnd_mat, curand_rand_mat.cu:30-53) ------------------
// z(i, j) = sqrt(-2 ln u1) cos(2 pi u2), (u1, u2) from Philox counter (i, j, kStreamOmega+index, 0).
__global__ void gaussian_fill_kernel(uint64_t seed, uint32_t stream, int64_t n, int32_t r, float* out, int32_t ld) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * r) return;
  const int64_t i = t / r;
  const int32_t j = (int32_t)(t - i * r);
  const u32x4 w = draw(seed, stream, (uint32_t)i, (uint32_t)j, 0);
  const double u1 = u53_open0(w.x, w.y);
  const double u2 = u53(w.z, w.w);
  const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
  out[i * ld + j] = (float)z;
}

int gaussian_fill(hipStream_t s, uint64_t seed, uint32_t index, int64_t n, int32_t r, float* out, int32_t ld) {
  MMSB_CHECK_ARG(n >= 0 && r >= 1 && ld >= r && n < ((int64_t)1 << 32), "gaussian: bad shape");
  if (n == 0) return MMSB_OK;
  const int64_t total = n * r;
  hipLaunchKernelGGL(gaussian_fill_kernel, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, s, seed,
                     (uint32_t)kStreamOmega + index, n, r, out, ld);
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

// ---- column means from the transposed CSR: out[j] = scale * sum_{e in row j of G^T} val ------
__global__ void csr_rowsum_kernel(int32_t m, const int32_t* __restrict__ rowptr, const float* __restrict__ vals,
                                  double scale, double* __restrict__ out) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double acc;
  if (vals) {
    acc = 0.0;
    for (int32_t e = rowptr[i]; e < rowptr[i + 1]; ++e) acc += (double)vals[e];
  } else {
    acc = (double)(rowptr[i + 1] - rowptr[i]);
  }
  out[i] = scale * acc;
}

int csr_rowsum(hipStream_t s, int32_t m, const int32_t* rowptr, const float* vals, double scale, double* out) {
  if (m <= 0) return MMSB_OK;
  hipLaunchKernelGGL(csr_rowsum_kernel, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, s, m, rowptr, vals, scale, out);
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

// ---- C[p x q] (FP64) = alpha * A^T B, A [n x p], B [n x q] row-major ------------------------
// Grid: (tiles of C 64 x 64) x S row slices. Each workgroup stages 16-row slabs of A and B through
// LDS as FP64 and accumulates a 4 x 4 register tile per thread; the S partial tiles are written to
// a workspace and reduced in a fixed order by a second kernel (deterministic result).
template <typename TA, typename TB>
__global__ __launch_bounds__(256) void tn_partial_kernel(int64_t n, int32_t p, int32_t q, const TA* __restrict__ A,
                                                         int32_t lda, const TB* __restrict__ B, int32_t ldb,
                                                         int32_t slices, double* __restrict__ work) {
  __shared__ double As[16][64 + 1];
  __shared__ double Bs[16][64 + 1];
  const int tp = blockIdx.x * 64, tq = blockIdx.y * 64;
  const int slice = blockIdx.z;
  const int64_t r0 = n * slice / slices, r1 = n * (slice + 1) / slices;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;  // output rows tp+4tx.., cols tq+4ty..
  double acc[4][4] = {};
  for (int64_t base = r0; base < r1; base += 16) {
    // 16 x 64 slab of A and of B: 1024 elements each, 4 per thread
    for (int e = threadIdx.x; e < 16 * 64; e += 256) {
      const int rr = e >> 6, cc = e & 63;
      const int64_t row = base + rr;
      As[rr][cc] = (row < r1 && tp + cc < p) ? (double)A[row * lda + tp + cc] : 0.0;
      Bs[rr][cc] = (row < r1 && tq + cc < q) ? (double)B[row * ldb + tq + cc] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int rr = 0; rr < 16; ++rr) {
      double a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = As[rr][4 * tx + u];
        b[u] = Bs[rr][4 * ty + u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = fma(a[u], b[v], acc[u][v]);
    }
    __syncthreads();
  }
  double* w = work + (int64_t)slice * p * q;
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = tp + 4 * tx + u, j = tq + 4 * ty + v;
      if (i < p && j < q) w[(int64_t)i * q + j] = acc[u][v];
    }
}

__global__ void tn_reduce_kernel(int64_t pq, int32_t slices, const double* __restrict__ work, double alpha,
                                 double beta, double* __restrict__ C) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= pq) return;
  double acc = 0.0;
  for (int s = 0; s < slices; ++s) acc += work[(int64_t)s * pq + t];
  C[t] = (beta == 0.0) ? alpha * acc : alpha * acc + beta * C[t];
}

// ---- the vector shapes over many rows: C[1 x m] = v^T M (p = 1) or C[m x 1] = M^T v (q = 1) ----------
// One thread per output column walks a slice of the rows (M's rows read coalesced, v[i] the same for
// the whole wave) with four independent FMA chains; tn_reduce_kernel then sums the slices in a fixed
// order. (The tile kernel below would stage 64-column slabs of which one column is used.)
constexpr int64_t kVecMinRows = 65536;

static int tn_vec_slices(int64_t n, int32_t m) {
  const int64_t blocks = ceil_div(m, 256);
  int64_t s = ceil_div(2048, blocks);                // aim for ~2048 workgroups
  s = std::min<int64_t>(s, std::max<int64_t>(1, n / 256));
  return (int)std::max<int64_t>(1, std::min<int64_t>(s, 1024));
}

template <typename TM, typename TV>
__global__ __launch_bounds__(256) void tn_vec_kernel(int64_t n, int32_t m, const TM* __restrict__ M, int32_t ldm,
                                                     const TV* __restrict__ v, int32_t incv, int32_t slices,
                                                     double* __restrict__ work) {
  const int32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= m) return;
  const int64_t r0 = n * blockIdx.y / slices, r1 = n * (blockIdx.y + 1) / slices;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int64_t i = r0;
  for (; i + 4 <= r1; i += 4) {
    a0 = fma((double)v[i * incv], (double)M[i * ldm + j], a0);
    a1 = fma((double)v[(i + 1) * incv], (double)M[(i + 1) * ldm + j], a1);
    a2 = fma((double)v[(i + 2) * incv], (double)M[(i + 2) * ldm + j], a2);
    a3 = fma((double)v[(i + 3) * incv], (double)M[(i + 3) * ldm + j], a3);
  }
  for (; i < r1; ++i) a0 = fma((double)v[i * incv], (double)M[i * ldm + j], a0);
  work[(int64_t)blockIdx.y * m + j] = (a0 + a1) + (a2 + a3);
}

static int tn_slices(int64_t n, int32_t p, int32_t q) {
  const int64_t tiles = ceil_div(p, 64) * ceil_div(q, 64);
  int64_t s = ceil_div(2048, tiles);                 // aim for ~2048 workgroups
  s = std::min<int64_t>(s, std::max<int64_t>(1, n / 256));
  return (int)std::max<int64_t>(1, std::min<int64_t>(s, 512));
}

// ---- FP64 MFMA tall products ---------------------------------------------------------------------
// v_mfma_f64_16x16x4f64: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16] (one double each);
// accumulator register v of lane l holds D[4 v + l / 16][l % 16].
// Workgroup tile 128 x 128 of the output, four waves of 64 x 64 (4 x 4 MFMA tiles each); the
// reduction runs in stages of 16 through LDS (one buffer, the next stage's global loads in flight
// in registers during the current stage's MFMAs). Inputs are converted to FP64 on the way into LDS,
// so FP32 data gives the same products as the FP64 vector kernels.
constexpr int kMT = 128;        // workgroup tile
constexpr int kMK = 16;         // reduction stage
constexpr int kMPad = kMT + 16; // LDS row stride (doubles) of the k-major stages

static bool use_mfma(int64_t n, int32_t p, int32_t q) { return n >= 4096 && p >= 32 && q >= 32; }

static int tn_mfma_tiles(int32_t p, int32_t q, bool sym) {
  const int tp = (int)ceil_div(p, kMT), tq = (int)ceil_div(q, kMT);
  return sym ? tp * (tp + 1) / 2 : tp * tq;
}

// Slices of the reduction: about 1024 workgroups in all, i.e. two full rounds of the 512 that fit at
// once (256 CUs x 2); rounding the count down keeps the last round full (10 symmetric tiles of a
// 510 x 510 Gram: 102 slices, where 103 capped to 64 left the second round a quarter full).
static int tn_mfma_slices(int64_t n, int32_t p, int32_t q, bool sym) {
  const int64_t tiles = tn_mfma_tiles(p, q, sym);
  int64_t s = std::max<int64_t>(1, 1024 / tiles);
  const int64_t cap = std::max<int64_t>(1, (int64_t)(256ll << 20) / ((int64_t)p * q * 8));  // partials <= 256 MiB
  s = std::min<int64_t>(std::min<int64_t>(s, cap), std::max<int64_t>(1, n / 1024));
  return (int)std::max<int64_t>(1, s);
}

size_t tn_workspace_bytes(int64_t n, int32_t p, int32_t q) {
  size_t b = sizeof(double) * (size_t)tn_slices(n, p, q) * (size_t)p * (size_t)q;
  if ((p == 1 || q == 1) && n >= kVecMinRows)  // bounded by the slice cap, so that a narrower call fits too
    b = std::max(b, sizeof(double) * (size_t)std::min<int64_t>(1024, std::max<int64_t>(1, n / 256)) * (size_t)p *
                        (size_t)q);
  if (use_mfma(n, p, q))
    for (bool sym : {false, true})
      b = std::max(b, sizeof(double) * (size_t)tn_mfma_slices(n, p, q, sym) * (size_t)p * (size_t)q);
  return b;
}

__device__ inline doublex4 mfma_f64(double a, double b, doublex4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// One wave's 64 x 64 block from k-major LDS stages: As[k][i] (rows of the output), Bs[k][j].
__device__ inline void mfma_stage(doublex4 (&acc)[4][4], const double (*As)[kMPad], const double (*Bs)[kMPad], int ai,
                                  int bj) {
  const int l = threadIdx.x & 63, lr = l & 15, lk = l >> 4;
#pragma unroll
  for (int kk = 0; kk < kMK; kk += 4) {
    double a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a[u] = As[kk + lk][ai + 16 * u + lr];
      b[u] = Bs[kk + lk][bj + 16 * u + lr];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[u][v] = mfma_f64(a[u], b[v], acc[u][v]);
  }
}

// C partial [slice] = A[r0:r1]^T B[r0:r1] for output tile (ti, tj); SYM: A == B, tiles ti <= tj only.
template <typename TA, typename TB, bool SYM>
__global__ __launch_bounds__(256) void tn_mfma_kernel(int64_t n, int32_t p, int32_t q, const TA* __restrict__ A,
                                                      int32_t lda, const TB* __restrict__ B, int32_t ldb, int32_t slices,
                                                      double* __restrict__ work) {
  __shared__ double As[kMK][kMPad];
  __shared__ double Bs[kMK][kMPad];
  int ti, tj;
  if constexpr (SYM) {  // upper-triangular tile pairs in row order
    const int tp = (int)((p + kMT - 1) / kMT);
    int rem = blockIdx.x;
    ti = 0;
    while (rem >= tp - ti) {
      rem -= tp - ti;
      ++ti;
    }
    tj = ti + rem;
  } else {
    const int tq = (int)((q + kMT - 1) / kMT);
    ti = blockIdx.x / tq;
    tj = blockIdx.x % tq;
  }
  const int slice = blockIdx.y;
  const int64_t r0 = n * slice / slices, r1 = n * (slice + 1) / slices;
  const int i0 = ti * kMT, j0 = tj * kMT;
  const int t = threadIdx.x, w = t >> 6;
  const int ai = 64 * (w >> 1), bj = 64 * (w & 1);
  // staging map: thread -> stage row sr, columns sc + 16 e
  const int sr = t >> 4, sc = t & 15;
  double ra[8], rb[8];
  auto load = [&](int64_t base) {
    const int64_t row = base + sr;
    const bool rv = row < r1;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int ci = i0 + sc + 16 * e, cj = j0 + sc + 16 * e;
      ra[e] = (rv && ci < p) ? (double)A[row * lda + ci] : 0.0;
      rb[e] = (rv && cj < q) ? (double)B[row * ldb + cj] : 0.0;
    }
  };
  doublex4 acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[u][v] = doublex4{0.0, 0.0, 0.0, 0.0};
  load(r0);
  for (int64_t base = r0; base < r1; base += kMK) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      As[sr][sc + 16 * e] = ra[e];
      Bs[sr][sc + 16 * e] = rb[e];
    }
    __syncthreads();
    if (base + kMK < r1) load(base + kMK);
    mfma_stage(acc, As, Bs, ai, bj);
  }
  double* wp = work + (int64_t)slice * p * q;
  const int l = t & 63, lr = l & 15, lk = l >> 4;
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int i = i0 + ai + 16 * u + 4 * g + lk, j = j0 + bj + 16 * v + lr;
        if (i < p && j < q) wp[(int64_t)i * q + j] = acc[u][v][g];
      }
}

// C = alpha * sum_s work[s] + beta * C; SYM: entries below the diagonal tiles come from the mirror.
__global__ void tn_mfma_reduce_kernel(int32_t p, int32_t q, int32_t slices, bool sym, const double* __restrict__ work,
                                      double alpha, double beta, double* __restrict__ C) {
  const int64_t pq = (int64_t)p * q;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= pq) return;
  const int i = (int)(t / q), j = (int)(t % q);
  const int64_t src = (sym && i / kMT > j / kMT) ? (int64_t)j * q + i : t;
  double acc = 0.0;
  for (int s = 0; s < slices; ++s) acc += work[(int64_t)s * pq + src];
  C[t] = (beta == 0.0) ? alpha * acc : alpha * acc + beta * C[t];
}

template <typename TA, typename TB>
int gemm_tn(hipStream_t s, int64_t n, int32_t p, int32_t q, double alpha, const TA* A, int32_t lda, const TB* B,
            int32_t ldb, double beta, double* C, double* work) {
  MMSB_CHECK_ARG(n >= 0 && p >= 1 && q >= 1 && lda >= p && ldb >= q, "gemm_tn: bad shape");
  const int64_t pq = (int64_t)p * q;
  if ((p == 1 || q == 1) && n >= kVecMinRows) {
    const int32_t m = p == 1 ? q : p;
    const int slices = tn_vec_slices(n, m);
    const dim3 grid((unsigned)ceil_div(m, 256), (unsigned)slices);
    if (p == 1)  // C = A^T B with A a column: v = A, M = B
      hipLaunchKernelGGL((tn_vec_kernel<TB, TA>), grid, dim3(256), 0, s, n, m, B, ldb, A, lda, slices, work);
    else         // q == 1: v = B, M = A
      hipLaunchKernelGGL((tn_vec_kernel<TA, TB>), grid, dim3(256), 0, s, n, m, A, lda, B, ldb, slices, work);
    MMSB_LAUNCH_CHECK();
    hipLaunchKernelGGL(tn_reduce_kernel, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, s, (int64_t)m, slices, work,
                       alpha, beta, C);
    MMSB_LAUNCH_CHECK();
    return MMSB_OK;
  }
  if (use_mfma(n, p, q)) {
    const bool sym = std::is_same<TA, TB>::value && (const void*)A == (const void*)B && lda == ldb && p == q;
    const int slices = tn_mfma_slices(n, p, q, sym);
    dim3 grid((unsigned)tn_mfma_tiles(p, q, sym), (unsigned)slices);
    if (sym)
      hipLaunchKernelGGL((tn_mfma_kernel<TA, TB, true>), grid, dim3(256), 0, s, n, p, q, A, lda, B, ldb, slices, work);
    else
      hipLaunchKernelGGL((tn_mfma_kernel<TA, TB, false>), grid, dim3(256), 0, s, n, p, q, A, lda, B, ldb, slices, work);
    MMSB_LAUNCH_CHECK();
    hipLaunchKernelGGL(tn_mfma_reduce_kernel, dim3((unsigned)ceil_div(pq, 256)), dim3(256), 0, s, p, q, slices, sym, work,
                       alpha, beta, C);
    MMSB_LAUNCH_CHECK();
    return MMSB_OK;
  }
  const int slices = tn_slices(n, p, q);
  dim3 grid((unsigned)ceil_div(p, 64), (unsigned)ceil_div(q, 64), (unsigned)slices);
  hipLaunchKernelGGL((tn_partial_kernel<TA, TB>), grid, dim3(256), 0, s, n, p, q, A, lda, B, ldb, slices, work);
  MMSB_LAUNCH_CHECK();
  hipLaunchKernelGGL(tn_reduce_kernel, dim3((unsigned)ceil_div(pq, 256)), dim3(256), 0, s, pq, slices, work, alpha,
                     beta, C);
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

template int gemm_tn<float, float>(hipStream_t, int64_t, int32_t, int32_t, double, const float*, int32_t, const float*,
                                   int32_t, double, double*, double*);
template int gemm_tn<float, double>(hipStream_t, int64_t, int32_t, int32_t, double, const float*, int32_t,
                                    const double*, int32_t, double, double*, double*);
template int gemm_tn<double, double>(hipStream_t, int64_t, int32_t, int32_t, double, const double*, int32_t,
                                     const double*, int32_t, double, double*, double*);
template int gemm_tn<double, float>(hipStream_t, int64_t, int32_t, int32_t, double, const double*, int32_t,
                                    const float*, int32_t, double, double*, double*);

// ---- Y [n x q] = alpha * X [n x p] * C [p x q] (+ beta * Y), C FP64 ---------------------------
// Workgroup: 64 rows x 64 columns of Y; the p-loop stages 16-wide slabs of X and C through LDS.
template <typename TX, typename TY, bool CT>
__global__ __launch_bounds__(256) void nn_kernel(int64_t n, int32_t p, int32_t q, double alpha, const TX* __restrict__ X,
                                                 int32_t ldx, const double* __restrict__ C, int32_t ldc, double beta,
                                                 TY* __restrict__ Y, int32_t ldy) {
  __shared__ double Xs[64][16 + 1];
  __shared__ double Cs[16][64 + 1];
  const int64_t r0 = (int64_t)blockIdx.x * 64;
  const int c0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;  // rows r0 + ty*4.., cols c0 + tx*4..
  double acc[4][4] = {};
  for (int kb = 0; kb < p; kb += 16) {
    for (int e = threadIdx.x; e < 64 * 16; e += 256) {
      const int rr = e >> 4, kk = e & 15;
      const int64_t row = r0 + rr;
      Xs[rr][kk] = (row < n && kb + kk < p) ? (double)X[row * ldx + kb + kk] : 0.0;
      const int kk2 = e >> 6, cc = e & 63;
      const int64_t cidx = CT ? (int64_t)(c0 + cc) * ldc + kb + kk2 : (int64_t)(kb + kk2) * ldc + c0 + cc;
      Cs[kk2][cc] = (kb + kk2 < p && c0 + cc < q) ? C[cidx] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < 16; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = Xs[4 * ty + u][kk];
        b[u] = Cs[kk][4 * tx + u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = fma(a[u], b[v], acc[u][v]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int64_t row = r0 + 4 * ty + u;
    if (row >= n) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int col = c0 + 4 * tx + v;
      if (col < q) {
        TY* y = Y + row * ldy + col;
        const double o = alpha * acc[u][v];
        *y = (TY)((beta == 0.0) ? o : o + beta * (double)*y);
      }
    }
  }
}

// Y tile [128 x 128] = alpha * X[rows, :] C[:, cols] (+ beta * Y) on FP64 MFMA; CT: C stored transposed.
template <typename TX, typename TY, bool CT>
__global__ __launch_bounds__(256, 2) void nn_mfma_kernel(int64_t n, int32_t p, int32_t q, double alpha,
                                                      const TX* __restrict__ X, int32_t ldx, const double* __restrict__ C,
                                                      int32_t ldc, double beta, TY* __restrict__ Y, int32_t ldy) {
  __shared__ double Xs[kMK][kMPad];  // k-major: Xs[k][row]
  __shared__ double Cs[kMK][kMPad];  // Cs[k][col]
  const int64_t r0 = (int64_t)blockIdx.x * kMT;
  const int c0 = blockIdx.y * kMT;
  const int t = threadIdx.x, w = t >> 6;
  const int ai = 64 * (w >> 1), bj = 64 * (w & 1);
  // X stage map: thread -> row xr + 32 e' ... : 128 rows x 16 k = 2048 values, 8 per thread
  const int xk = t & 15, xr = t >> 4;  // k = xk, rows xr + 16 e
  double rx[8], rc[8];
  auto load = [&](int kb) {
    const bool kv = kb + xk < p;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int64_t row = r0 + xr + 16 * e;
      rx[e] = (kv && row < n) ? (double)X[row * ldx + kb + xk] : 0.0;
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      if constexpr (CT) {  // C[col][k]: thread -> k = t & 15, cols (t >> 4) + 16 e
        const int col = c0 + xr + 16 * e;
        rc[e] = (kv && col < q) ? C[(int64_t)col * ldc + kb + xk] : 0.0;
      } else {             // C[k][col]: thread -> k = t >> 4, cols (t & 15) + 16 e
        const int k = kb + xr, col = c0 + xk + 16 * e;
        rc[e] = (k < p && col < q) ? C[(int64_t)k * ldc + col] : 0.0;
      }
    }
  };
  doublex4 acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[u][v] = doublex4{0.0, 0.0, 0.0, 0.0};
  load(0);
  for (int kb = 0; kb < p; kb += kMK) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      Xs[xk][xr + 16 * e] = rx[e];
      if constexpr (CT)
        Cs[xk][xr + 16 * e] = rc[e];
      else
        Cs[xr][xk + 16 * e] = rc[e];
    }
    __syncthreads();
    if (kb + kMK < p) load(kb + kMK);
    mfma_stage(acc, Xs, Cs, ai, bj);
  }
  const int l = t & 63, lr = l & 15, lk = l >> 4;
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int64_t row = r0 + ai + 16 * u + 4 * g + lk;
      if (row >= n) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int col = c0 + bj + 16 * v + lr;
        if (col < q) {
          TY* y = Y + row * ldy + col;
          const double o = alpha * acc[u][v][g];
          *y = (TY)((beta == 0.0) ? o : o + beta * (double)*y);
        }
      }
    }
}

template <typename TX, typename TY>
int gemm_nn(hipStream_t s, int64_t n, int32_t p, int32_t q, double alpha, const TX* X, int32_t ldx, const double* C,
            int32_t ldc, double beta, TY* Y, int32_t ldy, bool c_trans) {
  MMSB_CHECK_ARG(n >= 0 && p >= 1 && q >= 1 && ldx >= p && ldy >= q && ldc >= (c_trans ? p : q), "gemm_nn: bad shape");
  if (n == 0) return MMSB_OK;
  if (use_mfma(n, p, q)) {
    dim3 g((unsigned)ceil_div(n, kMT), (unsigned)ceil_div(q, kMT));
    if (c_trans)
      hipLaunchKernelGGL((nn_mfma_kernel<TX, TY, true>), g, dim3(256), 0, s, n, p, q, alpha, X, ldx, C, ldc, beta, Y, ldy);
    else
      hipLaunchKernelGGL((nn_mfma_kernel<TX, TY, false>), g, dim3(256), 0, s, n, p, q, alpha, X, ldx, C, ldc, beta, Y, ldy);
    MMSB_LAUNCH_CHECK();
    return MMSB_OK;
  }
  dim3 grid((unsigned)ceil_div(n, 64), (unsigned)ceil_div(q, 64));
  if (c_trans)
    hipLaunchKernelGGL((nn_kernel<TX, TY, true>), grid, dim3(256), 0, s, n, p, q, alpha, X, ldx, C, ldc, beta, Y, ldy);
  else
    hipLaunchKernelGGL((nn_kernel<TX, TY, false>), grid, dim3(256), 0, s, n, p, q, alpha, X, ldx, C, ldc, beta, Y, ldy);
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

template int gemm_nn<float, float>(hipStream_t, int64_t, int32_t, int32_t, double, const float*, int32_t,
                                   const double*, int32_t, double, float*, int32_t, bool);
template int gemm_nn<double, double>(hipStream_t, int64_t, int32_t, int32_t, double, const double*, int32_t,
                                     const double*, int32_t, double, double*, int32_t, bool);
template int gemm_nn<float, double>(hipStream_t, int64_t, int32_t, int32_t, double, const float*, int32_t,
                                    const double*, int32_t, double, double*, int32_t, bool);
template int gemm_nn<double, float>(hipStream_t, int64_t, int32_t, int32_t, double, const double*, int32_t,
                                    const double*, int32_t, double, float*, int32_t, bool);

// ---- FP32-MFMA tall product for well-conditioned small factors -----------------------------------
// Y [n x q] = X [n x p] C [p x q] with X, Y FP32 and C FP64 rounded to FP32 (CT: C given as [q x p]),
// on v_mfma_f32_32x32x2f32 with FP32 sums: 128 x 128 workgroup tiles, four waves of 64 x 64 (2 x 2
// MFMA blocks; lane l supplies A[l % 32][l / 32] and B[l / 32][l % 32], accumulator register i of
// lane l holds D[8 (i / 4) + 4 (l / 32) + i % 4][l % 32]), 16-deep k stages through LDS. Meant for
// the second CholeskyQR pass (C = R2^{-1}, R2 = I + O(kappa eps)) and products with orthogonal r x r
// factors, where the FP32 rounding of C and the FP32 sums (~sqrt(p) eps_32 relative) stay at the
// FP32 storage level of Y.
constexpr int kFT = 128, kFK = 16, kFPad = kFT + 4;

template <bool CT>
__global__ __launch_bounds__(256, 2) void nn_f32mfma_kernel(int64_t n, int32_t p, int32_t q, float alpha,
                                                           const float* __restrict__ X, int32_t ldx,
                                                           const double* __restrict__ C, int32_t ldc, float beta,
                                                           float* __restrict__ Y, int32_t ldy) {
  __shared__ float Xs[kFK][kFPad];  // Xs[k][row]
  __shared__ float Cs[kFK][kFPad];  // Cs[k][col]
  const int64_t r0 = (int64_t)blockIdx.x * kFT;
  const int c0 = blockIdx.y * kFT;
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  const int ai = 64 * (w >> 1), bj = 64 * (w & 1);
  const int xk = t & 15, xr = t >> 4;  // stage map: k = xk, rows xr + 16 e
  float rx[8], rc[8];
  auto load = [&](int kb) {
    const bool kv = kb + xk < p;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int64_t row = r0 + xr + 16 * e;
      rx[e] = (kv && row < n) ? X[row * ldx + kb + xk] : 0.f;
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      if constexpr (CT) {  // C[col][k]: k = xk, cols xr + 16 e
        const int col = c0 + xr + 16 * e;
        rc[e] = (kv && col < q) ? (float)C[(int64_t)col * ldc + kb + xk] : 0.f;
      } else {             // C[k][col]: k = xr, cols xk + 16 e
        const int k = kb + xr, col = c0 + xk + 16 * e;
        rc[e] = (k < p && col < q) ? (float)C[(int64_t)k * ldc + col] : 0.f;
      }
    }
  };
  floatx16 acc[2][2];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v) acc[u][v] = floatx16{};
  load(0);
  for (int kb = 0; kb < p; kb += kFK) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      Xs[xk][xr + 16 * e] = rx[e];
      if constexpr (CT)
        Cs[xk][xr + 16 * e] = rc[e];
      else
        Cs[xr][xk + 16 * e] = rc[e];
    }
    __syncthreads();
    if (kb + kFK < p) load(kb + kFK);
#pragma unroll
    for (int kk = 0; kk < kFK; kk += 2) {
      const int kl = kk + (l >> 5);
      float a[2], b[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        a[u] = Xs[kl][ai + 32 * u + (l & 31)];
        b[u] = Cs[kl][bj + 32 * u + (l & 31)];
      }
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v) acc[u][v] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[u], b[v], acc[u][v], 0, 0, 0);
    }
  }
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v)
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int64_t row = r0 + ai + 32 * u + 8 * (i >> 2) + 4 * (l >> 5) + (i & 3);
        const int col = c0 + bj + 32 * v + (l & 31);
        if (row < n && col < q) {
          float* y = Y + row * ldy + col;
          *y = beta == 0.f ? alpha * acc[u][v][i] : alpha * acc[u][v][i] + beta * *y;
        }
      }
}

int gemm_nn_f32(hipStream_t s, int64_t n, int32_t p, int32_t q, double alpha, const float* X, int32_t ldx,
                const double* C, int32_t ldc, double beta, float* Y, int32_t ldy, bool c_trans) {
  MMSB_CHECK_ARG(n >= 0 && p >= 1 && q >= 1 && ldx >= p && ldy >= q && ldc >= (c_trans ? p : q), "gemm_nn_f32: bad shape");
  if (n == 0) return MMSB_OK;
  const dim3 g((unsigned)ceil_div(n, kFT), (unsigned)ceil_div(q, kFT));
  const float a = (float)alpha, b = (float)beta;
  if (c_trans)
    hipLaunchKernelGGL(nn_f32mfma_kernel<true>, g, dim3(256), 0, s, n, p, q, a, X, ldx, C, ldc, b, Y, ldy);
  else
    hipLaunchKernelGGL(nn_f32mfma_kernel<false>, g, dim3(256), 0, s, n, p, q, a, X, ldx, C, ldc, b, Y, ldy);
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

// C [p x q] (FP64) = alpha A^T B + beta C with A [n x p], B [n x q] FP32: the rows are cut into the
// slices of tn_mfma_slices; each workgroup sums its slice of one 128 x 128 tile on
// v_mfma_f32_32x32x2f32 (FP32 sums over at most n / slices rows), the partials are written as FP64
// and reduced in FP64 by tn_mfma_reduce_kernel. For products whose relative accuracy need not beat
// ~sqrt(n / slices) eps_32 (the operator applications of the randomized range finder).
__global__ __launch_bounds__(256, 2) void tn_f32mfma_kernel(int64_t n, int32_t p, int32_t q, const float* __restrict__ A,
                                                           int32_t lda, const float* __restrict__ B, int32_t ldb,
                                                           int32_t slices, double* __restrict__ work) {
  __shared__ float As[kFK][kFPad];  // As[row][i]
  __shared__ float Bs[kFK][kFPad];  // Bs[row][j]
  const int tq = (int)((q + kFT - 1) / kFT);
  const int ti = blockIdx.x / tq, tj = blockIdx.x % tq;
  const int slice = blockIdx.y;
  const int64_t r0 = n * slice / slices, r1 = n * (slice + 1) / slices;
  const int i0 = ti * kFT, j0 = tj * kFT;
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  const int ai = 64 * (w >> 1), bj = 64 * (w & 1);
  const int sr = t >> 4, sc = t & 15;  // stage map: row sr, columns sc + 16 e
  float ra[8], rb[8];
  auto load = [&](int64_t base) {
    const int64_t row = base + sr;
    const bool rv = row < r1;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int ci = i0 + sc + 16 * e, cj = j0 + sc + 16 * e;
      ra[e] = (rv && ci < p) ? A[row * lda + ci] : 0.f;
      rb[e] = (rv && cj < q) ? B[row * ldb + cj] : 0.f;
    }
  };
  floatx16 acc[2][2];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v) acc[u][v] = floatx16{};
  load(r0);
  for (int64_t base = r0; base < r1; base += kFK) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      As[sr][sc + 16 * e] = ra[e];
      Bs[sr][sc + 16 * e] = rb[e];
    }
    __syncthreads();
    if (base + kFK < r1) load(base + kFK);
#pragma unroll
    for (int kk = 0; kk < kFK; kk += 2) {
      const int kl = kk + (l >> 5);
      float a[2], b[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        a[u] = As[kl][ai + 32 * u + (l & 31)];
        b[u] = Bs[kl][bj + 32 * u + (l & 31)];
      }
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v) acc[u][v] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[u], b[v], acc[u][v], 0, 0, 0);
    }
  }
  double* wp = work + (int64_t)slice * p * q;
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v)
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int ri = i0 + ai + 32 * u + 8 * (i >> 2) + 4 * (l >> 5) + (i & 3);
        const int cj = j0 + bj + 32 * v + (l & 31);
        if (ri < p && cj < q) wp[(int64_t)ri * q + cj] = (double)acc[u][v][i];
      }
}

int gemm_tn_f32(hipStream_t s, int64_t n, int32_t p, int32_t q, double alpha, const float* A, int32_t lda, const float* B,
                int32_t ldb, double beta, double* C, double* work) {
  MMSB_CHECK_ARG(n >= 0 && p >= 1 && q >= 1 && lda >= p && ldb >= q, "gemm_tn_f32: bad shape");
  const int slices = tn_mfma_slices(n, p, q, false);
  const dim3 grid((unsigned)(ceil_div(p, kFT) * ceil_div(q, kFT)), (unsigned)slices);
  hipLaunchKernelGGL(tn_f32mfma_kernel, grid, dim3(256), 0, s, n, p, q, A, lda, B, ldb, slices, work);
  MMSB_LAUNCH_CHECK();
  const int64_t pq = (int64_t)p * q;
  hipLaunchKernelGGL(tn_mfma_reduce_kernel, dim3((unsigned)ceil_div(pq, 256)), dim3(256), 0, s, p, q, slices, false, work,
                     alpha, beta, C);
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

// ---- Cholesky factor and its inverse (CholeskyQR of the tall whitening sketches) ----------------
// The r x r FP64 matrices stay in L2. Right-looking blocked Cholesky of the lower triangle of
// G + shift trace(G) I in 16-column panels, two launches per panel: chol_panel_kernel (one
// workgroup) factors the diagonal block in LDS and solves the rows below it, one row per thread;
// chol_trail_kernel updates the trailing lower triangle in 64 x 64 tiles spread over workgroups.
// R = L^T and R^{-1} = X^T with X = L^{-1}: for r <= 1024, chol_transpose_kernel writes R through
// LDS tiles and trinv_cols_kernel solves blocks of kTrinvCols columns of X in parallel workgroups
// (one thread per row of X); larger r falls back to trinv_kernel, one workgroup that walks PI rows
// at a time with one thread per column of X.
constexpr int kCholP = 16;
constexpr int kCholThreads = 512;

// Diagonal shift G += max(shift trace(G), 1e-300) I; clears the status word.
__global__ __launch_bounds__(kCholThreads) void chol_shift_kernel(int32_t r, double* __restrict__ G, double shift,
                                                                  int32_t* __restrict__ status) {
  __shared__ double red[kCholThreads];
  const int t = threadIdx.x;
  double tr = 0.0;
  for (int i = t; i < r; i += blockDim.x) tr += G[(int64_t)i * r + i];
  red[t] = tr;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (t < off) red[t] += red[t + off];
    __syncthreads();
  }
  double sh = shift * red[0];
  if (!(sh > 0.0)) sh = 1e-300;
  for (int i = t; i < r; i += blockDim.x) G[(int64_t)i * r + i] += sh;
  if (t == 0) status[0] = 0;
}

// Panel j0 (columns j0 .. j0 + pb): the diagonal block factored in LDS, then the rows below it solved
// against it, L21 = A21 L11^{-T}, one row per thread. A non-positive pivot is replaced by 1 and the
// first such column (1-based) is recorded in status[0].
__global__ __launch_bounds__(kCholThreads) void chol_panel_kernel(int32_t r, double* __restrict__ G, int32_t j0,
                                                                  int32_t* __restrict__ status) {
  __shared__ double D[kCholP][kCholP + 1];
  __shared__ int bad;
  const int t = threadIdx.x;
  const int pb = min(kCholP, r - j0);
  if (t == 0) bad = 0;
  if (t < kCholP * kCholP) {
    const int a = t / kCholP, b = t % kCholP;
    if (a < pb && b <= a) D[a][b] = G[(int64_t)(j0 + a) * r + j0 + b];
  }
  __syncthreads();
  for (int c = 0; c < pb; ++c) {
    if (t == 0) {
      const double d = D[c][c];
      if (!(d > 0.0)) {
        if (!bad) bad = j0 + c + 1;
        D[c][c] = 1.0;
      } else {
        D[c][c] = sqrt(d);
      }
    }
    __syncthreads();
    if (t > c && t < pb) D[t][c] /= D[c][c];
    __syncthreads();
    if (t < kCholP * kCholP) {
      const int a = t / kCholP, b = t % kCholP;
      if (a < pb && b > c && b <= a) D[a][b] -= D[a][c] * D[b][c];
    }
    __syncthreads();
  }
  if (t < kCholP * kCholP) {
    const int a = t / kCholP, b = t % kCholP;
    if (a < pb && b <= a) G[(int64_t)(j0 + a) * r + j0 + b] = D[a][b];
  }
  for (int i = j0 + pb + t; i < r; i += blockDim.x) {
    __asm__ volatile("" ::: "memory");  // re-read D per row instead of holding 136 values in registers
    double* row = G + (int64_t)i * r + j0;
    double x[kCholP];
#pragma unroll
    for (int c = 0; c < kCholP; ++c) x[c] = c < pb ? row[c] : 0.0;
#pragma unroll
    for (int c = 0; c < kCholP; ++c) {
      if (c < pb) {
        double v = x[c];
#pragma unroll
        for (int cc = 0; cc < c; ++cc) v -= x[cc] * D[c][cc];
        x[c] = v / D[c][c];
        row[c] = x[c];
      }
    }
  }
  if (t == 0 && bad && status[0] == 0) status[0] = bad;
}

// Trailing update of panel j0: G[i][l] -= sum_c L[i][c] L[l][c] over the lower triangle
// j0 + pb <= l <= i < r, one 64 x 64 tile pair (ti >= tj) per workgroup, the panel rows of both tiles
// staged in LDS, 4 x 4 outputs per thread (the same summation order as a single-workgroup update).
constexpr int kTrailT = 64;

__global__ __launch_bounds__(256) void chol_trail_kernel(int32_t r, double* __restrict__ G, int32_t j0) {
  __shared__ double Ls[2][kTrailT][kCholP + 1];
  const int pb = min(kCholP, r - j0);
  const int base = j0 + pb;
  int rem = blockIdx.x, ti = 0;
  while (rem > ti) {
    rem -= ti + 1;
    ++ti;
  }
  const int tj = rem;
  const int i0 = base + ti * kTrailT, l0 = base + tj * kTrailT;
  const int t = threadIdx.x;
  for (int e = t; e < kTrailT * kCholP; e += blockDim.x) {
    const int rr = e / kCholP, c = e % kCholP;
    const int i = i0 + rr, l = l0 + rr;
    Ls[0][rr][c] = (i < r && c < pb) ? G[(int64_t)i * r + j0 + c] : 0.0;
    Ls[1][rr][c] = (l < r && c < pb) ? G[(int64_t)l * r + j0 + c] : 0.0;
  }
  __syncthreads();
  const int bu = 4 * (t >> 4), bv = 4 * (t & 15);
  double acc[4][4] = {};
  for (int c = 0; c < pb; ++c) {
    double a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a[u] = Ls[0][bu + u][c];
      b[u] = Ls[1][bv + u][c];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[u][v] = fma(a[u], b[v], acc[u][v]);
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = i0 + bu + u, l = l0 + bv + v;
      if (i < r && l <= i) G[(int64_t)i * r + l] -= acc[u][v];
    }
}

template <int PI>
__global__ __launch_bounds__(1024) void trinv_kernel(int32_t r, const double* __restrict__ L, double* __restrict__ R,
                                                     double* __restrict__ Rinv) {
  __shared__ double Lr[16384];  // PI rows of L, r entries each
  const int t = threadIdx.x;
  const int64_t nel = (int64_t)r * r;
  for (int64_t e = t; e < nel; e += blockDim.x) {
    const int i = (int)(e / r), l = (int)(e % r);
    R[e] = (l >= i) ? L[(int64_t)l * r + i] : 0.0;  // R = L^T
    Rinv[e] = 0.0;
  }
  for (int i0 = 0; i0 < r; i0 += PI) {
    const int pb = min(PI, r - i0);
    __syncthreads();  // previous panel's Rinv entries and Lr reads are complete
    for (int e = t; e < pb * r; e += blockDim.x) {
      const int ii = e / r, l = e % r;
      Lr[ii * r + l] = (l <= i0 + ii) ? L[(int64_t)(i0 + ii) * r + l] : 0.0;
    }
    __syncthreads();
    for (int c = t; c < i0 + pb; c += blockDim.x) {
      double acc[PI];
#pragma unroll
      for (int ii = 0; ii < PI; ++ii) acc[ii] = 0.0;
      const double* xc = Rinv + (int64_t)c * r;  // row c of R^{-1} = column c of X
      for (int l = c; l < i0; ++l) {
        const double x = xc[l];
#pragma unroll
        for (int ii = 0; ii < PI; ++ii) acc[ii] = fma(Lr[ii * r + l], x, acc[ii]);
      }
      double xs[PI];
#pragma unroll
      for (int ii = 0; ii < PI; ++ii) {
        const int i = i0 + ii;
        double v = 0.0;
        if (ii < pb && i >= c) {
          v = (i == c ? 1.0 : 0.0) - acc[ii];
#pragma unroll
          for (int jj = 0; jj < ii; ++jj) v = fma(-Lr[ii * r + i0 + jj], xs[jj], v);
          v /= Lr[ii * r + i];
        }
        xs[ii] = v;
      }
#pragma unroll
      for (int ii = 0; ii < PI; ++ii)
        if (ii < pb) Rinv[(int64_t)c * r + i0 + ii] = xs[ii];
    }
  }
}

// R = L^T (upper triangle, zeros below) from the lower triangle of L, through a 32 x 32 LDS tile so
// that reads and writes are both along rows.
__global__ __launch_bounds__(256) void chol_transpose_kernel(int32_t r, const double* __restrict__ L,
                                                             double* __restrict__ R) {
  __shared__ double t[32][33];
  const int bi = blockIdx.y * 32, bj = blockIdx.x * 32;  // R rows bi.., R columns bj..
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int y = ty; y < 32; y += 8) {
    const int l = bj + y, i = bi + tx;  // L[l][i] = R[i][l]
    t[y][tx] = (l < r && i < r) ? L[(int64_t)l * r + i] : 0.0;
  }
  __syncthreads();
  for (int y = ty; y < 32; y += 8) {
    const int i = bi + y, l = bj + tx;
    if (i < r && l < r) R[(int64_t)i * r + l] = (l >= i) ? t[tx][y] : 0.0;
  }
}

// Columns [c0, c0 + kTrinvCols) of X = L^{-1} by forward substitution, one thread per row k of X
// (blockDim >= r): b_k starts as the unit vector entries; at step i the owner of row i finalises
// x_i = b_i / L_ii and broadcasts it through LDS (double-buffered by step parity, one barrier per
// step), and every row k > i subtracts L_ki x_i. L_ki = R[i][k] is read along row i of R (coalesced),
// kTrinvRing steps ahead of its use. Rows of X above a column's diagonal stay exactly zero. Writes
// row c of R^{-1} (= column c of X) for each of the block's columns.
constexpr int kTrinvCols = 16;
constexpr int kTrinvRing = 16;

__global__ __launch_bounds__(1024) void trinv_cols_kernel(int32_t r, const double* __restrict__ R,
                                                          double* __restrict__ Rinv) {
  constexpr int C = kTrinvCols, D = kTrinvRing;
  __shared__ double xb[2][C];
  const int k = threadIdx.x;
  const int c0 = blockIdx.x * C;
  double b[C];
#pragma unroll
  for (int c = 0; c < C; ++c) b[c] = (k == c0 + c) ? 1.0 : 0.0;
  const double inv_kk = k < r ? 1.0 / R[(int64_t)k * r + k] : 0.0;
  auto lval = [&](int i) -> double { return (i < r && k > i && k < r) ? R[(int64_t)i * r + k] : 0.0; };
  double ring[D];
#pragma unroll
  for (int d = 0; d < D; ++d) ring[d] = lval(c0 + d);
  for (int ib = c0; ib < r; ib += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int i = ib + d;
      if (i < r) {  // uniform over the block
        const double lki = ring[d];
        ring[d] = lval(i + D);
        if (k == i) {
#pragma unroll
          for (int c = 0; c < C; ++c) {
            b[c] *= inv_kk;
            xb[i & 1][c] = b[c];
          }
        }
        __syncthreads();
        if (k > i) {
#pragma unroll
          for (int c = 0; c < C; ++c) b[c] = fma(-lki, xb[i & 1][c], b[c]);
        }
      }
    }
  }
  if (k < r) {
#pragma unroll
    for (int c = 0; c < C; ++c)
      if (c0 + c < r) Rinv[(int64_t)(c0 + c) * r + k] = (k >= c0 + c) ? b[c] : 0.0;
  }
}

int chol_inv(hipStream_t s, int32_t r, double* G, double shift, double* R, double* Rinv, int32_t* status) {
  MMSB_CHECK_ARG(r >= 1 && r <= 4096, "chol_inv: r=%d out of range", r);
  hipLaunchKernelGGL(chol_shift_kernel, dim3(1), dim3(kCholThreads), 0, s, r, G, shift, status);
  MMSB_LAUNCH_CHECK();
  for (int j0 = 0; j0 < r; j0 += kCholP) {
    hipLaunchKernelGGL(chol_panel_kernel, dim3(1), dim3(kCholThreads), 0, s, r, G, j0, status);
    MMSB_LAUNCH_CHECK();
    const int m = r - std::min(r, j0 + kCholP);
    if (m > 0) {
      const int tt = (m + kTrailT - 1) / kTrailT;
      hipLaunchKernelGGL(chol_trail_kernel, dim3(tt * (tt + 1) / 2), dim3(256), 0, s, r, G, j0);
      MMSB_LAUNCH_CHECK();
    }
  }
  if (r <= 1024) {
    const int nt = (r + 31) / 32;
    hipLaunchKernelGGL(chol_transpose_kernel, dim3(nt, nt), dim3(256), 0, s, r, G, R);
    MMSB_LAUNCH_CHECK();
    const int threads = (r + 63) / 64 * 64;
    hipLaunchKernelGGL(trinv_cols_kernel, dim3((r + kTrinvCols - 1) / kTrinvCols), dim3(threads), 0, s, r, R, Rinv);
  } else if (r <= 2048)
    hipLaunchKernelGGL(trinv_kernel<8>, dim3(1), dim3(1024), 0, s, r, G, R, Rinv);
  else
    hipLaunchKernelGGL(trinv_kernel<4>, dim3(1), dim3(1024), 0, s, r, G, R, Rinv);
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

__global__ void fill_u64_kernel(int64_t n, unsigned long long* p, unsigned long long v) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

// ---- column sign canonicalisation: largest-|.| entry of each column made positive ----------
// key = (|v| bits << 32) | ~row: atomicMax picks the largest magnitude, then the smallest row.
template <typename T>
__global__ void colmax_kernel(int64_t n, int32_t c, const T* __restrict__ M, int32_t ld,
                              unsigned long long* __restrict__ best) {
  const int col = blockIdx.y * blockDim.x + threadIdx.x;
  if (col >= c) return;
  const int64_t r0 = n * blockIdx.x / gridDim.x, r1 = n * (blockIdx.x + 1) / gridDim.x;
  unsigned long long b = 0;
  for (int64_t i = r0; i < r1; ++i) {
    const float a = fabsf((float)M[i * ld + col]);
    const unsigned long long key = ((unsigned long long)__float_as_uint(a) << 32) | (0xFFFFFFFFull - (unsigned long long)i);
    b = key > b ? key : b;
  }
  atomicMax(best + col, b);
}

template <typename T>
__global__ void colsign_kernel(int32_t c, const T* __restrict__ M, int32_t ld, const unsigned long long* __restrict__ best,
                               double* __restrict__ sign) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= c) return;
  const int64_t row = (int64_t)(0xFFFFFFFFull - (best[col] & 0xFFFFFFFFull));
  sign[col] = ((double)M[row * ld + col] < 0.0) ? -1.0 : 1.0;
}

template <typename T>
__global__ void scale_cols_kernel(int64_t n, int32_t c, T* __restrict__ M, int32_t ld, const double* __restrict__ sc) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * c) return;
  const int64_t i = t / c;
  const int j = (int)(t - i * c);
  M[i * ld + j] = (T)((double)M[i * ld + j] * sc[j]);
}

template <typename T>
int col_sign(hipStream_t s, int64_t n, int32_t c, const T* M, int32_t ld, unsigned long long* best, double* sign) {
  MMSB_CHECK_ARG(n >= 1 && n < 0xFFFFFFFFll && c >= 1, "col_sign: bad shape");
  hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)ceil_div(c, 256)), dim3(256), 0, s, (int64_t)c, best, 0ull);
  MMSB_LAUNCH_CHECK();
  const int slabs = (int)std::min<int64_t>(256, ceil_div(n, 1024));
  dim3 grid((unsigned)slabs, (unsigned)ceil_div(c, 64));
  hipLaunchKernelGGL((colmax_kernel<T>), grid, dim3(64), 0, s, n, c, M, ld, best);
  MMSB_LAUNCH_CHECK();
  hipLaunchKernelGGL((colsign_kernel<T>), dim3((unsigned)ceil_div(c, 64)), dim3(64), 0, s, c, M, ld, best, sign);
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

template <typename T>
int scale_cols(hipStream_t s, int64_t n, int32_t c, T* M, int32_t ld, const double* sc) {
  if (n <= 0 || c <= 0) return MMSB_OK;
  hipLaunchKernelGGL((scale_cols_kernel<T>), dim3((unsigned)ceil_div(n * c, 256)), dim3(256), 0, s, n, c, M, ld, sc);
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

template int col_sign<float>(hipStream_t, int64_t, int32_t, const float*, int32_t, unsigned long long*, double*);
template int col_sign<double>(hipStream_t, int64_t, int32_t, const double*, int32_t, unsigned long long*, double*);
template int scale_cols<float>(hipStream_t, int64_t, int32_t, float*, int32_t, const double*);
template int scale_cols<double>(hipStream_t, int64_t, int32_t, double*, int32_t, const double*);

// ---- small helpers ---------------------------------------------------------------------------
// out[c][r] = in[r][c] for an R x C FP32 matrix (32 x 32 LDS tiles)
__global__ void transpose_kernel(int64_t R, int32_t C, const float* __restrict__ in, int64_t ldi, float* __restrict__ out,
                                 int64_t ldo) {
  __shared__ float t[32][33];
  const int64_t r0 = (int64_t)blockIdx.x * 32;
  const int c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 32 x 8
  for (int j = ty; j < 32; j += 8) {
    const int64_t r = r0 + j;
    const int c = c0 + tx;
    t[j][tx] = (r < R && c < C) ? in[r * ldi + c] : 0.f;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int c = c0 + j;
    const int64_t r = r0 + tx;
    if (c < C && r < R) out[(int64_t)c * ldo + r] = t[tx][j];
  }
}

int transpose_f32(hipStream_t s, int64_t R, int32_t C, const float* in, int64_t ldi, float* out, int64_t ldo) {
  if (R <= 0 || C <= 0) return MMSB_OK;
  hipLaunchKernelGGL(transpose_kernel, dim3((unsigned)ceil_div(R, 32), (unsigned)ceil_div(C, 32)), dim3(256), 0, s, R, C, in,
                     ldi, out, ldo);
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

// sinv[j] = sig[j] > tol * sig[0] ? 1 / sig[j] : 0      (pinv_vector, CPU Spectral.cpp:50-57)
// mode 1: out[j] = lam[j] > tol * lam[0] ? 1 / sqrt(lam[j]) : 0 (W = U Lambda^-1/2, CPU Util.cpp:119)
__global__ void pinv_scale_kernel(int32_t k, const double* __restrict__ v, double shift, double tol, int mode,
                                  double* __restrict__ out, double* __restrict__ vout) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k) return;
  const double x = v[j] - shift, x0 = v[0] - shift;
  if (vout) vout[j] = x;
  const bool keep = x > tol * (x0 > 0.0 ? x0 : 0.0) && x > 0.0;
  out[j] = keep ? (mode == 0 ? 1.0 / x : 1.0 / sqrt(x)) : 0.0;
}

int pinv_scale(hipStream_t s, int32_t k, const double* v, double shift, double tol, int mode, double* out, double* vout) {
  hipLaunchKernelGGL(pinv_scale_kernel, dim3((unsigned)ceil_div(k, 256)), dim3(256), 0, s, k, v, shift, tol, mode, out, vout);
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

// S <- (S + S^T)/2 + c I with c = ||S||_F (makes the Rayleigh-Ritz matrix positive semidefinite
// so that its SVD is its eigendecomposition); c written to shift_out. One workgroup.
__global__ __launch_bounds__(1024) void sym_shift_kernel(int32_t r, double* __restrict__ S, double* __restrict__ shift_out) {
  __shared__ double red[1024];
  const int t = threadIdx.x;
  const int64_t rr = (int64_t)r * r;
  double acc = 0.0;
  for (int64_t e = t; e < rr; e += blockDim.x) acc += S[e] * S[e];
  red[t] = acc;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (t < off) red[t] += red[t + off];
    __syncthreads();
  }
  const double c = sqrt(red[0]);
  for (int64_t e = t; e < rr; e += blockDim.x) {
    const int i = (int)(e / r), j = (int)(e % r);
    if (i < j) {
      const double v = 0.5 * (S[e] + S[(int64_t)j * r + i]);
      S[e] = v;
      S[(int64_t)j * r + i] = v;
    }
  }
  __syncthreads();
  for (int i = t; i < r; i += blockDim.x) S[(int64_t)i * r + i] += c;
  if (t == 0) shift_out[0] = c;
}

int sym_shift(hipStream_t s, int32_t r, double* S, double* shift_out) {
  hipLaunchKernelGGL(sym_shift_kernel, dim3(1), dim3(1024), 0, s, r, S, shift_out);
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

// P[i][j] *= alpha * d[i]  (k x c FP64 row-major)
__global__ void scale_rows_kernel(int32_t k, int32_t c, double* __restrict__ P, int32_t ld, const double* __restrict__ d,
                                  double alpha) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)k * c) return;
  const int i = (int)(t / c), j = (int)(t % c);
  P[(int64_t)i * ld + j] *= alpha * d[i];
}

int scale_rows(hipStream_t s, int32_t k, int32_t c, double* P, int32_t ld, const double* d, double alpha) {
  hipLaunchKernelGGL(scale_rows_kernel, dim3((unsigned)ceil_div((int64_t)k * c, 256)), dim3(256), 0, s, k, c, P, ld, d, alpha);
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

// OUT[i][j] += alpha * mu[i]^2 * X[i][j]   (the diagonal of mu mu^T added back, GPU M2 scaling)
__global__ void diag_axpy_kernel(int64_t n, int32_t c, const double* __restrict__ mu, double alpha, const float* __restrict__ X,
                                 int32_t ldx, float* __restrict__ O, int32_t ldo) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * c) return;
  const int64_t i = t / c;
  const int j = (int)(t - i * c);
  O[i * ldo + j] = (float)((double)O[i * ldo + j] + alpha * mu[i] * mu[i] * (double)X[i * ldx + j]);
}

int diag_axpy(hipStream_t s, int64_t n, int32_t c, const double* mu, double alpha, const float* X, int32_t ldx, float* O,
              int32_t ldo) {
  if (n <= 0) return MMSB_OK;
  hipLaunchKernelGGL(diag_axpy_kernel, dim3((unsigned)ceil_div(n * c, 256)), dim3(256), 0, s, n, c, mu, alpha, X, ldx, O, ldo);
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

// copy a k x k FP32 block into a zero-padded kp x kp buffer (and back)
__global__ void pad_copy_kernel(int32_t rows, int32_t cols, const float* __restrict__ in, int32_t ldi, float* __restrict__ out,
                                int32_t ldo) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)rows * cols) return;
  const int i = (int)(t / cols), j = (int)(t % cols);
  out[(int64_t)i * ldo + j] = in[(int64_t)i * ldi + j];
}

int copy2d_f32(hipStream_t s, int32_t rows, int32_t cols, const float* in, int32_t ldi, float* out, int32_t ldo) {
  if (rows <= 0 || cols <= 0) return MMSB_OK;
  hipLaunchKernelGGL(pad_copy_kernel, dim3((unsigned)ceil_div((int64_t)rows * cols, 256)), dim3(256), 0, s, rows, cols, in, ldi,
                     out, ldo);
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

}  // namespace mmsb
