// STGD step in BF16 MFMA with FP32 accumulation (same mathematics as stgd.hip).
//
// Data contractions -- P_A = Y_A[X_b] Phi, P_B = Y_B[X_b] Phi and Y_C[X_b]^T C_g -- take BF16
// operands (Y stored in BF16, Phi and C_g rounded to BF16) on v_mfma_f32_32x32x16_bf16 with FP32
// accumulation. The k x k x k products (Phi^T Phi, Phi (G o G) and the Newton-Schulz iterations)
// use split BF16: x = hi + lo with hi = bf16(x), lo = bf16(x - hi), and
// x.y ~ hi.hi + hi.lo + lo.hi, which keeps ~16 significant bits (FP32 inputs, FP32 accumulation) at
// three BF16 MFMA passes. Every GEMM operand here is K-contiguous (A[row][k] and B^T[col][k]), so
// Phi is kept in both layouts (Phi and Phi^T, FP32) and C_g is written transposed.
// An FP32 polishing pass (stgd.hip) normally follows the BF16 epochs.
#include "common.h"
#include "stgd.h"

namespace mmsb {

static constexpr int KCB = 32;        // K slab (elements)
static constexpr int PB = KCB + 8;    // LDS row pitch in bf16 (80 bytes)

__device__ inline __bf16 f2bf(float x) { return (__bf16)x; }  // round to nearest even

// Stage a 64 x KCB slab of a K-contiguous operand into LDS (hi, and lo when SPLIT).
template <typename T, bool SPLIT>
__device__ inline void stage_kc(__bf16 (*Hs)[PB], __bf16 (*Ls)[PB], const T* __restrict__ X, int64_t ld, int64_t r0,
                                int64_t k0) {
  const int t = threadIdx.x;
  if constexpr (sizeof(T) == 2) {
    // 64 rows x 32 bf16 = 64 B per row = 4 x 16 B; 256 threads, one 16-byte load each
    const int rr = t >> 2, q = t & 3;
    const uint4 v = *reinterpret_cast<const uint4*>(X + (r0 + rr) * ld + k0 + 8 * q);
    *reinterpret_cast<uint4*>(&Hs[rr][8 * q]) = v;
  } else {
    // 64 rows x 32 floats = 128 B per row = 8 x float4; two float4 per thread
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      const int e = t + 256 * it;
      const int rr = e >> 3, q = e & 7;
      const float4 v = *reinterpret_cast<const float4*>(X + (r0 + rr) * ld + k0 + 4 * q);
      const float f[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const __bf16 h = f2bf(f[u]);
        Hs[rr][4 * q + u] = h;
        if (SPLIT) Ls[rr][4 * q + u] = f2bf(f[u] - (float)h);
      }
    }
  }
}

// acc (wave's 32x32 block) += A[r0.., :K] * Bt[c0.., :K]^T
template <typename TA, typename TB, bool SPLIT>
__device__ inline void tile_bf16(floatx16& acc, __bf16 (*Ah)[PB], __bf16 (*Al)[PB], __bf16 (*Bh)[PB], __bf16 (*Bl)[PB],
                                 const TA* __restrict__ A, int64_t lda, const TB* __restrict__ Bt, int64_t ldb, int64_t r0,
                                 int64_t c0, int64_t K) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int fr = lane & 31, fk = 8 * (lane >> 5);
  for (int64_t k0 = 0; k0 < K; k0 += KCB) {
    stage_kc<TA, SPLIT>(Ah, Al, A, lda, r0, k0);
    stage_kc<TB, SPLIT>(Bh, Bl, Bt, ldb, c0, k0);
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KCB; kk += 16) {
      const bf16x8 ah = *reinterpret_cast<const bf16x8*>(&Ah[wr + fr][kk + fk]);
      const bf16x8 bh = *reinterpret_cast<const bf16x8*>(&Bh[wc + fr][kk + fk]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
      if (SPLIT) {
        const bf16x8 al = *reinterpret_cast<const bf16x8*>(&Al[wr + fr][kk + fk]);
        const bf16x8 bl = *reinterpret_cast<const bf16x8*>(&Bl[wc + fr][kk + fk]);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
      }
    }
    __syncthreads();
  }
}

__device__ inline int acc_row16(int i, int lane) { return 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3); }

#define BF16_LDS                         \
  __shared__ __bf16 Ah[64][PB];          \
  __shared__ __bf16 Al[64][PB];          \
  __shared__ __bf16 Bh[64][PB];          \
  __shared__ __bf16 Bl[64][PB]

// ---- phase 1: C_g tiles (BF16) and H = (Phi^T Phi)^o2 tiles (split BF16) -----------------------
__global__ __launch_bounds__(256) void bf_phase1_kernel(StgdArgs a, const __bf16* __restrict__ YA,
                                                       const __bf16* __restrict__ YB, const float* __restrict__ phiT,
                                                       __bf16* __restrict__ cgt16, int64_t row0, int32_t nb) {
  BF16_LDS;
  __shared__ float mvec[2][64];
  __shared__ float csum[4][32];
  const int kp = a.kp, ntk = kp / 64, ntb = a.b / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  int tile = blockIdx.x;
  if (tile < ntb * ntk) {
    const int tb = tile / ntk, tj = tile % ntk;
    const int64_t x0 = (int64_t)tb * 64, j0 = (int64_t)tj * 64;
    if (threadIdx.x < 128) {  // m_A, m_B = Phi^T mu for this tile's columns (FP32)
      const int which = threadIdx.x >> 6, c = threadIdx.x & 63;
      const double* mu = which ? a.mu_b : a.mu_a;
      const float* col = phiT + (j0 + c) * (int64_t)kp;
      float s = 0.f;
      for (int m = 0; m < kp; ++m) s = fmaf((float)mu[m], col[m], s);
      mvec[which][c] = s;
    }
    floatx16 pa = {}, pb = {};
    tile_bf16<__bf16, float, false>(pa, Ah, Al, Bh, Bl, YA + row0 * (int64_t)kp, kp, phiT, kp, x0, j0, kp);
    tile_bf16<__bf16, float, false>(pb, Ah, Al, Bh, Bl, YB + row0 * (int64_t)kp, kp, phiT, kp, x0, j0, kp);
    const int col = lane & 31;
    const float mA = mvec[0][wc + col], mB = mvec[1][wc + col];
    float cs = 0.f;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int row = wr + acc_row16(i, lane);
      const float p = pa[i], q = pb[i];
      const float ab = p * q;
      const bool valid = (x0 + row) < nb;
      cs += valid ? ab : 0.f;
      const float cg = valid ? a.c3 * ab + a.c1 * (p * mB + mA * q) : 0.f;
      cgt16[(j0 + wc + col) * (int64_t)a.b + x0 + row] = f2bf(cg);
    }
    cs += __shfl_xor(cs, 32);
    if (lane < 32) csum[wave][lane] = cs;
    __syncthreads();
    if (threadIdx.x < 64) {
      const int c = threadIdx.x, wcol = c >> 5;
      a.cs_part[(int64_t)tb * kp + j0 + c] = csum[wcol][c & 31] + csum[wcol + 2][c & 31];
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
    tile_bf16<float, float, true>(g, Ah, Al, Bh, Bl, phiT, kp, phiT, kp, i0, j0, kp);  // G = Phi^T Phi
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int row = wr + acc_row16(i, lane), col = wc + (lane & 31);
      a.h[(i0 + row) * kp + j0 + col] = g[i] * g[i];
    }
  }
}

// ---- phase 2: Phi~ (and Phi~^T) ----------------------------------------------------------------
__global__ __launch_bounds__(256) void bf_phase2_kernel(StgdArgs a, const __bf16* __restrict__ YCt,
                                                       const __bf16* __restrict__ cgt16, float* __restrict__ phitT,
                                                       int64_t row0, int32_t nb) {
  BF16_LDS;
  __shared__ float cmu[64];
  const int kp = a.kp, ntk = kp / 64;
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
  tile_bf16<float, float, true>(ph, Ah, Al, Bh, Bl, a.phi, kp, a.h, kp, i0, j0, kp);        // Phi H (H symmetric)
  tile_bf16<__bf16, __bf16, false>(t, Ah, Al, Bh, Bl, YCt + row0, a.n_pad, cgt16, a.b, i0, j0, a.b);  // Y_C^T C_g
  __syncthreads();
  const float eta = a.eta, inv_b = 1.f / (float)nb;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int row = wr + acc_row16(i, lane), col = wc + (lane & 31);
    const int64_t off = (i0 + row) * kp + j0 + col;
    const float muc = (float)a.mu_c[i0 + row];
    const float v = a.phi[off] - eta * a.theta * ph[i] + eta * (t[i] * inv_b + muc * cmu[col]);
    a.phit[off] = v;
    phitT[(j0 + col) * (int64_t)kp + i0 + row] = v;
  }
}

// ---- Newton-Schulz (split BF16): G = X^T X from X^T; X_out = 1.5 X - 0.5 X G ----------------------
__global__ __launch_bounds__(256) void bf_ns_gram_kernel(int32_t kp, const float* __restrict__ XT, float* __restrict__ G) {
  BF16_LDS;
  const int ntk = kp / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int ti = blockIdx.x / ntk, tj = blockIdx.x % ntk;
  const int64_t i0 = (int64_t)ti * 64, j0 = (int64_t)tj * 64;
  floatx16 g = {};
  tile_bf16<float, float, true>(g, Ah, Al, Bh, Bl, XT, kp, XT, kp, i0, j0, kp);
#pragma unroll
  for (int i = 0; i < 16; ++i) G[(i0 + wr + acc_row16(i, lane)) * kp + j0 + wc + (lane & 31)] = g[i];
}

__global__ __launch_bounds__(256) void bf_ns_update_kernel(int32_t kp, const float* __restrict__ X, const float* __restrict__ G,
                                                           float* __restrict__ Xout, float* __restrict__ XoutT,
                                                           const float* __restrict__ prev, double* __restrict__ frob) {
  BF16_LDS;
  __shared__ double red[4];
  const int ntk = kp / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int ti = blockIdx.x / ntk, tj = blockIdx.x % ntk;
  const int64_t i0 = (int64_t)ti * 64, j0 = (int64_t)tj * 64;
  floatx16 xg = {};
  tile_bf16<float, float, true>(xg, Ah, Al, Bh, Bl, X, kp, G, kp, i0, j0, kp);  // G symmetric
  double d = 0.0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int row = wr + acc_row16(i, lane), col = wc + (lane & 31);
    const int64_t off = (i0 + row) * kp + j0 + col;
    const float v = 1.5f * X[off] - 0.5f * xg[i];
    Xout[off] = v;
    XoutT[(j0 + col) * (int64_t)kp + i0 + row] = v;
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

__global__ void bf_copy_frob_kernel(int32_t kp, const float* __restrict__ X, float* __restrict__ Xout, float* __restrict__ XoutT,
                                    const float* __restrict__ prev, double* __restrict__ frob) {
  __shared__ double red[4];
  const int64_t nel = (int64_t)kp * kp;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double d = 0.0;
  if (t < nel) {
    const float v = X[t];
    Xout[t] = v;
    XoutT[(t % kp) * kp + t / kp] = v;
    const double dd = (double)v - (double)prev[t];
    d = dd * dd;
  }
  for (int off = 32; off > 0; off >>= 1) d += __shfl_xor(d, off);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = d;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(frob, red[0] + red[1] + red[2] + red[3]);
}

__global__ void bf_zero_f64(int64_t n, double* p) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.0;
}

__global__ void f32_to_bf16_kernel(int64_t n, const float* __restrict__ in, __bf16* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = f2bf(in[t]);
}

__global__ void transpose_sq_kernel(int32_t kp, const float* __restrict__ in, float* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < (int64_t)kp * kp) out[(t % kp) * kp + t / kp] = in[t];
}

int f32_to_bf16(hipStream_t s, int64_t n, const float* in, bf16_bits* out) {
  if (n <= 0) return MMSB_OK;
  hipLaunchKernelGGL(f32_to_bf16_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, n, in,
                     reinterpret_cast<__bf16*>(out));
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

int stgd_run_bf16(hipStream_t s, const StgdArgs& args, const bf16_bits* YA_, const bf16_bits* YB_, const bf16_bits* YCt_,
                  int64_t step0, int64_t nsteps, int32_t renorm_every, int32_t ns_iters, float* phi_buf[2], int& cur,
                  float* phit, float* gt, float* phia, float* phiT, float* phitT, float* phiaT, bf16_bits* cgt16_,
                  double* frob_log) {
  const __bf16* YA = reinterpret_cast<const __bf16*>(YA_);
  const __bf16* YB = reinterpret_cast<const __bf16*>(YB_);
  const __bf16* YCt = reinterpret_cast<const __bf16*>(YCt_);
  __bf16* cgt16 = reinterpret_cast<__bf16*>(cgt16_);
  MMSB_CHECK_ARG(args.kp % 64 == 0 && args.b % 64 == 0 && args.kp >= 64 && args.b >= 64, "stgd: kp and b must be multiples of 64");
  MMSB_CHECK_ARG(args.bl >= 1 && args.bl <= args.b, "stgd: logical minibatch %d outside [1, %d]", args.bl, args.b);
  MMSB_CHECK_ARG(args.bl % 8 == 0, "stgd bf16: the minibatch must be a multiple of 8 (got %d)", args.bl);
  MMSB_CHECK_ARG(args.n_pad % 8 == 0, "stgd bf16: n_pad must be a multiple of 8");
  const int64_t spe = ceil_div(args.n_x, args.bl);
  MMSB_CHECK_ARG(args.n_pad >= (spe - 1) * args.bl + args.b, "stgd: n_pad too small");
  const int kp = args.kp, ntk = kp / 64, ntb = args.b / 64;
  const int64_t nel = (int64_t)kp * kp;
  hipLaunchKernelGGL(bf_zero_f64, dim3((unsigned)ceil_div(nsteps, 256)), dim3(256), 0, s, nsteps, frob_log);
  MMSB_LAUNCH_CHECK();
  // Phi^T of the current Phi
  hipLaunchKernelGGL(transpose_sq_kernel, dim3((unsigned)ceil_div(nel, 256)), dim3(256), 0, s, kp, phi_buf[cur], phiT);
  MMSB_LAUNCH_CHECK();
  StgdArgs a = args;
  for (int64_t st = 0; st < nsteps; ++st) {
    const int64_t gstep = step0 + st;
    const int64_t row0 = (gstep % spe) * a.bl;
    const int32_t nb = (int32_t)std::min<int64_t>(a.bl, a.n_x - row0);
    a.phi = phi_buf[cur];
    a.phit = phit;
    hipLaunchKernelGGL(bf_phase1_kernel, dim3(ntb * ntk + ntk * ntk), dim3(256), 0, s, a, YA, YB, phiT, cgt16, row0, nb);
    MMSB_LAUNCH_CHECK();
    hipLaunchKernelGGL(bf_phase2_kernel, dim3(ntk * ntk), dim3(256), 0, s, a, YCt, cgt16, phitT, row0, nb);
    MMSB_LAUNCH_CHECK();
    float* out = phi_buf[cur ^ 1];
    double* frob = frob_log + st;
    const bool renorm = renorm_every > 0 && ((gstep + 1) % renorm_every == 0) && ns_iters > 0;
    if (renorm) {
      const float* x = phit;
      const float* xT = phitT;
      for (int it = 0; it < ns_iters; ++it) {
        const bool last = it == ns_iters - 1;
        float* dst = last ? out : (x == phia ? phit : phia);
        float* dstT = last ? phiT : (x == phia ? phitT : phiaT);
        hipLaunchKernelGGL(bf_ns_gram_kernel, dim3(ntk * ntk), dim3(256), 0, s, kp, xT, gt);
        MMSB_LAUNCH_CHECK();
        hipLaunchKernelGGL(bf_ns_update_kernel, dim3(ntk * ntk), dim3(256), 0, s, kp, x, gt, dst, dstT,
                           last ? a.phi : nullptr, frob);
        MMSB_LAUNCH_CHECK();
        x = dst;
        xT = dstT;
      }
    } else {
      hipLaunchKernelGGL(bf_copy_frob_kernel, dim3((unsigned)ceil_div(nel, 256)), dim3(256), 0, s, kp, phit, out, phiT, a.phi, frob);
      MMSB_LAUNCH_CHECK();
    }
    cur ^= 1;
  }
  return MMSB_OK;
}

}  // namespace mmsb
