// Stochastic tensor gradient descent on Phi (the hot loop of GPU_Task, stpm.c:497-790).
//
// Reference step (minibatch of one node, stpm.c:499-786):
//   phi~ = [mu_A, mu_B, y_A, y_B]^T Phi                       (stpm.c:532, 4 x k)
//   c_g  = phi~^T S_g phi~,  c_mu = phi~^T S_mu phi~           (stpm.c:547-595, 4k gemv launches)
//   T^   = y_C c_g^T + mu_C c_mu^T                            (stpm.c:624,635)
//   Phi <- Phi - eta Phi (Phi^T Phi)^o2 + eta T^              (stpm.c:649-691)
//   Phi <- QR(Phi)                                             (stpm.c:774-786)
// Here a step processes a minibatch X_b of b nodes at once:
//   P_A = Y_A[X_b] Phi, P_B = Y_B[X_b] Phi                    (b x k GEMMs)
//   C_g = c3 P_A o P_B + c1 (P_A o m_B + m_A o P_B)            (m_A = Phi^T mu_A, m_B = Phi^T mu_B)
//   c_mu = c0 m_A o m_B + c1 mean_x(P_A o P_B)
//   Phi~ = Phi - eta theta Phi (G o G) + eta [ Y_C[X_b]^T C_g / b + mu_C c_mu^T ],  G = Phi^T Phi
//   Phi <- polar(Phi~) by Newton-Schulz iterations Phi <- Phi (3I - Phi^T Phi)/2
// with (c3, c1, c0) = (1, -a0/(a0+2), 2a0^2/((a0+1)(a0+2))), the entries of the 4x4 matrices of
// init_buff (stpm.c:1026-1038). ||Phi_new - Phi||_F^2 of every step is logged on the device
// (the frob_norm record of stpm.c:723-744, without the per-step device-to-host copy).
//
// Layout: k is padded to kp, a multiple of 64; padded rows and columns of Phi, Y and mu are zero
// and stay exactly zero through every update, so every tile is full. Y_A, Y_B are [n_pad x kp]
// row-major (one node per row), Y_C is stored transposed [kp x n_pad] so that Y_C^T C_g reads
// contiguous node runs; C_g is written transposed [kp x b]. Phi is [kp x kp] row-major with
// column i the i-th eigenvector (Phi[m * kp + i]).
//
// GEMM tiles: a workgroup of 4 waves computes a 64 x 64 output tile, each wave a 32 x 32 block
// with v_mfma_f32_32x32x2f32 (FP32) or v_mfma_f32_32x32x16_bf16 (BF16 inputs, FP32 accumulate).
// Operand slabs of KC columns are staged through LDS in k-major layout; each operand is read from
// global memory along its contiguous dimension, whichever that is (template LoadMode).
#include "common.h"
#include "stgd.h"

namespace mmsb {

// ---- LDS-staged 64x64 FP32 MFMA tile ---------------------------------------------------------
// A(i, m): KCONTIG -> A[i * lda + m]; ROWCONTIG -> A[m * lda + i]
// B(m, j): KCONTIG -> B[j * ldb + m]; ROWCONTIG (contiguous in j) -> B[m * ldb + j]
enum LoadMode { KCONTIG = 0, ROWCONTIG = 1 };

static constexpr int KC = 32;         // K slab
static constexpr int LDSW = 64 + 1;   // padded LDS row (k-major: [KC][LDSW]); odd stride -> no bank conflicts

template <int MODE, typename T>
__device__ inline void stage_slab(float (*S)[LDSW], const T* __restrict__ X, int64_t ld, int64_t r0, int64_t k0) {
  // S[kk][rr] = X(r0 + rr, k0 + kk), rr < 64, kk < KC
  const int t = threadIdx.x;
  if (MODE == KCONTIG) {
#pragma unroll
    for (int e = t; e < 64 * KC; e += 256) {
      const int rr = e / KC, kk = e % KC;
      S[kk][rr] = (float)X[(r0 + rr) * ld + k0 + kk];
    }
  } else {
#pragma unroll
    for (int e = t; e < 64 * KC; e += 256) {
      const int kk = e / 64, rr = e % 64;
      S[kk][rr] = (float)X[(k0 + kk) * ld + r0 + rr];
    }
  }
}

// acc (wave's 32x32 block at (wr, wc) of the 64x64 tile) += A(r0.., :) B(:, c0..) over K.
template <int AMODE, int BMODE, typename TA, typename TB>
__device__ inline void tile_mma_f32(floatx16& acc, float (*As)[LDSW], float (*Bs)[LDSW], const TA* __restrict__ A,
                                    int64_t lda, const TB* __restrict__ B, int64_t ldb, int64_t r0, int64_t c0, int64_t K) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  for (int64_t k0 = 0; k0 < K; k0 += KC) {
    stage_slab<AMODE>(As, A, lda, r0, k0);
    stage_slab<BMODE>(Bs, B, ldb, c0, k0);
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KC; kk += 2) {
      const float a = As[kk + (lane >> 5)][wr + (lane & 31)];
      const float b = Bs[kk + (lane >> 5)][wc + (lane & 31)];
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
}

// element (row, col) of the wave's 32x32 block held in register i of this lane
__device__ inline int acc_row(int i, int lane) { return 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3); }
__device__ inline int acc_col(int lane) { return lane & 31; }

// ---- phase 1: P tiles (C_g) and G tiles (H) ---------------------------------------------------
// grid.x = ntile_b * ntile_k (P tiles) + ntile_k * ntile_k (G tiles)
template <typename TY>
__global__ __launch_bounds__(256) void stgd_phase1_kernel(StgdArgs a, const TY* __restrict__ YA, const TY* __restrict__ YB,
                                                         int64_t row0, int32_t nb) {
  __shared__ float As[KC][LDSW];
  __shared__ float Bs[KC][LDSW];
  __shared__ float mvec[2][64];
  __shared__ float csum[4][32];
  const int kp = a.kp;
  const int ntk = kp / 64;
  const int ntb = a.b / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const float* __restrict__ Phi = a.phi;
  int tile = blockIdx.x;
  if (tile < ntb * ntk) {
    const int tb = tile / ntk, tj = tile % ntk;
    const int64_t x0 = (int64_t)tb * 64, j0 = (int64_t)tj * 64;
    // m_A, m_B for the 64 columns of this tile: Phi^T mu
    if (threadIdx.x < 128) {
      const int which = threadIdx.x >> 6, c = threadIdx.x & 63;
      const double* mu = which ? a.mu_b : a.mu_a;
      float s = 0.f;
      for (int m = 0; m < kp; ++m) s = fmaf((float)mu[m], Phi[(int64_t)m * kp + j0 + c], s);
      mvec[which][c] = s;
    }
    floatx16 pa = {}, pb = {};
    tile_mma_f32<KCONTIG, ROWCONTIG>(pa, As, Bs, YA + row0 * (int64_t)kp, (int64_t)kp, Phi, (int64_t)kp, x0, j0, kp);
    tile_mma_f32<KCONTIG, ROWCONTIG>(pb, As, Bs, YB + row0 * (int64_t)kp, (int64_t)kp, Phi, (int64_t)kp, x0, j0, kp);
    // epilogue: C_g and the per-column sum of P_A o P_B over the (valid) rows of this tile
    const int col = acc_col(lane);
    const float mA = mvec[0][wc + col], mB = mvec[1][wc + col];
    float cs = 0.f;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int row = wr + acc_row(i, lane);
      const float p = pa[i], q = pb[i];
      const float ab = p * q;
      const bool valid = (x0 + row) < nb;
      cs += valid ? ab : 0.f;
      const float cg = valid ? a.c3 * ab + a.c1 * (p * mB + mA * q) : 0.f;
      a.cgt[(j0 + wc + col) * (int64_t)a.b + x0 + row] = cg;
    }
    // reduce cs over the two lane halves (rows) and the two row-waves
    cs += __shfl_xor(cs, 32);
    if (lane < 32) csum[wave][lane] = cs;
    __syncthreads();
    if (threadIdx.x < 64) {
      const int c = threadIdx.x;  // column within the tile
      const int wcol = c >> 5;    // waves 0/2 hold columns 0..31, waves 1/3 hold 32..63
      const float v = csum[wcol][c & 31] + csum[wcol + 2][c & 31];
      a.cs_part[(int64_t)tb * kp + j0 + c] = v;
      if (tb == 0) {
        a.mab[j0 + c] = mvec[0][c];
        a.mab[kp + j0 + c] = mvec[1][c];
      }
    }
  } else {
    tile -= ntb * ntk;
    const int ti = tile / ntk, tj = tile % ntk;
    const int64_t i0 = (int64_t)ti * 64, j0 = (int64_t)tj * 64;
    floatx16 g = {};
    // G(i, j) = sum_m Phi[m][i] Phi[m][j]
    tile_mma_f32<ROWCONTIG, ROWCONTIG>(g, As, Bs, Phi, (int64_t)kp, Phi, (int64_t)kp, i0, j0, kp);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int row = wr + acc_row(i, lane), col = wc + acc_col(lane);
      a.h[(i0 + row) * kp + j0 + col] = g[i] * g[i];
    }
  }
}

// ---- phase 2: Phi~ = Phi - eta theta Phi H + (eta/b) Y_C^T C_g + eta mu_C c_mu^T -------------
template <typename TY>
__global__ __launch_bounds__(256) void stgd_phase2_kernel(StgdArgs a, const TY* __restrict__ YCt, int64_t row0, int32_t nb) {
  __shared__ float As[KC][LDSW];
  __shared__ float Bs[KC][LDSW];
  __shared__ float cmu[64];
  const int kp = a.kp;
  const int ntk = kp / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int ti = blockIdx.x / ntk, tj = blockIdx.x % ntk;
  const int64_t i0 = (int64_t)ti * 64, j0 = (int64_t)tj * 64;
  if (threadIdx.x < 64) {
    const int c = threadIdx.x;
    float s = 0.f;
    for (int tb = 0; tb < a.b / 64; ++tb) s += a.cs_part[(int64_t)tb * kp + j0 + c];
    cmu[c] = a.c0 * a.mab[j0 + c] * a.mab[kp + j0 + c] + a.c1 * s / (float)nb;
  }
  floatx16 ph = {}, t = {};
  // Phi H: A = Phi (k-contiguous rows), B = H (contiguous in j)
  tile_mma_f32<KCONTIG, ROWCONTIG>(ph, As, Bs, a.phi, (int64_t)kp, a.h, (int64_t)kp, i0, j0, kp);
  // Y_C^T C_g: A(i, x) = YCt[i][row0 + x], B(x, j) = cgt[j][x]
  tile_mma_f32<KCONTIG, KCONTIG>(t, As, Bs, YCt + row0, (int64_t)a.n_pad, a.cgt, (int64_t)a.b, i0, j0, a.b);
  __syncthreads();
  const float eta = a.eta;
  const float inv_b = 1.f / (float)nb;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int row = wr + acc_row(i, lane), col = wc + acc_col(lane);
    const int64_t off = (i0 + row) * kp + j0 + col;
    const float muc = (float)a.mu_c[i0 + row];
    a.phit[off] = a.phi[off] - eta * a.theta * ph[i] + eta * (t[i] * inv_b + muc * cmu[col]);
  }
}

// ---- Newton-Schulz: G = X^T X, then X_out = 1.5 X - 0.5 X G (and the step's ||dPhi||^2) ------
__global__ __launch_bounds__(256) void ns_gram_kernel(int32_t kp, const float* __restrict__ X, float* __restrict__ G) {
  __shared__ float As[KC][LDSW];
  __shared__ float Bs[KC][LDSW];
  const int ntk = kp / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int ti = blockIdx.x / ntk, tj = blockIdx.x % ntk;
  const int64_t i0 = (int64_t)ti * 64, j0 = (int64_t)tj * 64;
  floatx16 g = {};
  tile_mma_f32<ROWCONTIG, ROWCONTIG>(g, As, Bs, X, (int64_t)kp, X, (int64_t)kp, i0, j0, kp);
#pragma unroll
  for (int i = 0; i < 16; ++i) G[(i0 + wr + acc_row(i, lane)) * kp + j0 + wc + acc_col(lane)] = g[i];
}

__global__ __launch_bounds__(256) void ns_update_kernel(int32_t kp, const float* __restrict__ X, const float* __restrict__ G,
                                                        float* __restrict__ Xout, const float* __restrict__ prev,
                                                        double* __restrict__ frob) {
  __shared__ float As[KC][LDSW];
  __shared__ float Bs[KC][LDSW];
  __shared__ double red[4];
  const int ntk = kp / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int ti = blockIdx.x / ntk, tj = blockIdx.x % ntk;
  const int64_t i0 = (int64_t)ti * 64, j0 = (int64_t)tj * 64;
  floatx16 xg = {};
  tile_mma_f32<KCONTIG, ROWCONTIG>(xg, As, Bs, X, (int64_t)kp, G, (int64_t)kp, i0, j0, kp);
  double d = 0.0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int64_t off = (i0 + wr + acc_row(i, lane)) * kp + j0 + wc + acc_col(lane);
    const float v = 1.5f * X[off] - 0.5f * xg[i];
    Xout[off] = v;
    if (prev) {
      const double dd = (double)v - (double)prev[off];
      d += dd * dd;
    }
  }
  if (prev) {
    for (int off = 32; off > 0; off >>= 1) d += __shfl_xor(d, off);
    if (lane == 0) red[wave] = d;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(frob, red[0] + red[1] + red[2] + red[3]);
  }
}

// Plain copy used when renormalisation is skipped for a step: Phi_next = Phi~ (+ ||dPhi||^2).
__global__ void copy_frob_kernel(int64_t nel, const float* __restrict__ X, float* __restrict__ Xout,
                                 const float* __restrict__ prev, double* __restrict__ frob) {
  __shared__ double red[4];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double d = 0.0;
  if (t < nel) {
    Xout[t] = X[t];
    const double dd = (double)X[t] - (double)prev[t];
    d = dd * dd;
  }
  for (int off = 32; off > 0; off >>= 1) d += __shfl_xor(d, off);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = d;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(frob, red[0] + red[1] + red[2] + red[3]);
}

__global__ void zero_f64_kernel(int64_t n, double* p) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.0;
}

// ---- host driver of one epoch --------------------------------------------------------------------
template <typename TY>
static int run_steps(hipStream_t s, StgdArgs a, const TY* YA, const TY* YB, const TY* YCt, int64_t step0, int64_t nsteps,
                     int64_t steps_per_epoch, int32_t renorm_every, int32_t ns_iters, float* phi_buf[2], int& cur,
                     float* phit, float* gt, float* phia, double* frob_log) {
  const int ntk = a.kp / 64;
  const int ntb = a.b / 64;
  for (int64_t st = 0; st < nsteps; ++st) {
    const int64_t gstep = step0 + st;
    const int64_t bi = gstep % steps_per_epoch;
    const int64_t row0 = bi * a.bl;
    const int32_t nb = (int32_t)std::min<int64_t>(a.bl, a.n_x - row0);
    a.phi = phi_buf[cur];
    a.phit = phit;
    hipLaunchKernelGGL((stgd_phase1_kernel<TY>), dim3(ntb * ntk + ntk * ntk), dim3(256), 0, s, a, YA, YB, row0, nb);
    MMSB_LAUNCH_CHECK();
    hipLaunchKernelGGL((stgd_phase2_kernel<TY>), dim3(ntk * ntk), dim3(256), 0, s, a, YCt, row0, nb);
    MMSB_LAUNCH_CHECK();
    float* out = phi_buf[cur ^ 1];
    double* frob = frob_log + st;
    const bool renorm = renorm_every > 0 && ((gstep + 1) % renorm_every == 0) && ns_iters > 0;
    if (renorm) {
      const float* x = phit;
      for (int it = 0; it < ns_iters; ++it) {
        const bool last = it == ns_iters - 1;
        float* dst = last ? out : (x == phia ? phit : phia);
        hipLaunchKernelGGL(ns_gram_kernel, dim3(ntk * ntk), dim3(256), 0, s, a.kp, x, gt);
        MMSB_LAUNCH_CHECK();
        hipLaunchKernelGGL(ns_update_kernel, dim3(ntk * ntk), dim3(256), 0, s, a.kp, x, gt, dst,
                           last ? a.phi : nullptr, frob);
        MMSB_LAUNCH_CHECK();
        x = dst;
      }
    } else {
      const int64_t nel = (int64_t)a.kp * a.kp;
      hipLaunchKernelGGL(copy_frob_kernel, dim3((unsigned)ceil_div(nel, 256)), dim3(256), 0, s, nel, phit, out, a.phi, frob);
      MMSB_LAUNCH_CHECK();
    }
    cur ^= 1;
  }
  return MMSB_OK;
}

int stgd_run_f32(hipStream_t s, const StgdArgs& args, const float* YA, const float* YB, const float* YCt, int64_t step0,
                 int64_t nsteps, int32_t renorm_every, int32_t ns_iters, float* phi_buf[2], int& cur, float* phit, float* gt,
                 float* phia, double* frob_log) {
  MMSB_CHECK_ARG(args.kp % 64 == 0 && args.b % 64 == 0 && args.kp >= 64 && args.b >= 64, "stgd: kp and b must be multiples of 64");
  MMSB_CHECK_ARG(args.bl >= 1 && args.bl <= args.b, "stgd: logical minibatch %d outside [1, %d]", args.bl, args.b);
  const int64_t spe = ceil_div(args.n_x, args.bl);
  MMSB_CHECK_ARG(args.n_pad >= (spe - 1) * args.bl + args.b, "stgd: n_pad too small");
  const int64_t nfrob = nsteps;
  hipLaunchKernelGGL(zero_f64_kernel, dim3((unsigned)ceil_div(nfrob, 256)), dim3(256), 0, s, nfrob, frob_log);
  MMSB_LAUNCH_CHECK();
  return run_steps<float>(s, args, YA, YB, YCt, step0, nsteps, spe, renorm_every, ns_iters, phi_buf, cur, phit, gt, phia, frob_log);
}

}  // namespace mmsb
