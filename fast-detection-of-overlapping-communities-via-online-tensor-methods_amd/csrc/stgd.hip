// STGD step kernels, FP32 and BF16 MFMA (the hot loop of GPU_Task, stpm.c:497-790).
//
// One step over a minibatch X_b of nb nodes (layouts in stgd.h):
//   P1  P_A = Y_A[X_b] Phi, P_B = Y_B[X_b] Phi, with m_A = Phi^T mu_A, m_B = Phi^T mu_B:
//       C_g = c3 P_A o P_B + c1 (P_A o m_B + m_A o P_B) (rows beyond nb zero);
//       per-tile column sums of P_A o P_B;  and  H = (Phi^T Phi) o (Phi^T Phi)
//   P2  Phi~ = Phi - eta theta Phi H + eta [Y_C[X_b]^T C_g / nb + mu_C c_mu^T],
//       c_mu = c0 m_A o m_B + c1 sum_x(P_A o P_B) / nb                       (stpm.c:532-691)
//   NS  G = X^T X; X <- c X (3I - c^2 G) / 2 with c = min(1, sqrt(2 / ||G||_inf)), ns_iters times
//       (polar renormalisation in place of the QR of stpm.c:774-786). The polar factor is
//       scale-invariant; c only acts when sigma_max(X)^2 may exceed 2, where the unscaled
//       iteration (convergent for sigma < sqrt 3) would diverge -- e.g. after a minibatch that
//       holds very high-degree nodes.
// (c3, c1, c0) are the entries of the 4x4 matrices of init_buff (stpm.c:1026-1038).
//
// Work decomposition: every GEMM output is cut into 32 x 32 blocks, one per workgroup of four
// waves. The waves split the reduction dimension and load their MFMA fragments straight from
// L2/HBM into registers (each wave owns its K range, so nothing is shared through LDS) through a
// ring of D fragment sets, so that up to D groups of loads are in flight ahead of the MFMAs. The
// four partial accumulators meet in a padded LDS tile that all 256 threads read back for
// coalesced epilogue stores. At k = 500 a phase launches 256-768 workgroups.
// K order inside a fragment: for v_mfma_f32_32x32x2f32 lane half h holds k = 4h + i of a chunk
// of eight k for instruction i (the order of a sum over k is free as long as A and B agree);
// for v_mfma_f32_32x32x16_bf16 lane half h holds the eight contiguous k = 8h .. 8h + 7.
// Precision: FP32 mode multiplies FP32 on v_mfma_f32_32x32x2f32. BF16 mode multiplies the data
// contractions (Y Phi, Y_C^T C_g) as BF16 on v_mfma_f32_32x32x16_bf16 and the k x k x k products
// as split BF16 (x = hi + lo; hi.hi + hi.lo + lo.hi, about 16 significant bits); all accumulate
// in FP32. In BF16 mode every k x k operand is written by the kernel that produces it both in
// FP32 and as its (hi, lo) BF16 pair, so the MFMA loops load ready operands and convert nothing.
// m_A and m_B of the next step are reduced from the tiles of the new Phi where it is written.
#include <type_traits>

#include "common.h"
#include "stgd.h"

namespace mmsb {

enum { MM_F32 = 0, MM_BF16 = 1, MM_BF16X3 = 2 };

template <int MODE>
using op_t = typename std::conditional<MODE == MM_F32, float, __bf16>::type;

// A matrix operand of wave_mm: rows of length >= K with leading dimension ld; lo is the low BF16
// half of a split operand (split mode only; equal to hi otherwise).
template <typename T>
struct Opnd {
  const T* hi;
  const T* lo;
  int64_t ld;
};

template <typename T>
__host__ __device__ inline Opnd<T> opnd(const T* hi, const T* lo, int64_t ld) {
  return Opnd<T>{hi, lo ? lo : hi, ld};
}

template <bool AL>
__device__ inline float4 ld4(const float* __restrict__ p) {
  if constexpr (AL) return *reinterpret_cast<const float4*>(p);
  return make_float4(p[0], p[1], p[2], p[3]);
}

template <bool AL>
__device__ inline bf16x8 ld8(const __bf16* __restrict__ p) {
  if constexpr (AL) return *reinterpret_cast<const bf16x8*>(p);
  bf16x8 v;
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = p[i];
  return v;
}

// Reduction group of one fragment set: 32 k for FP32, 64 k for the BF16 modes.
template <int MODE>
constexpr int kgroup() { return MODE == MM_F32 ? 32 : 64; }

// Fragment set of one reduction group for one wave.
template <int MODE>
struct Frag;

template <>
struct Frag<MM_F32> {
  float4 a[4], b[4];
  template <bool AAL>
  __device__ inline void load(const float* pa, const float*, const float* pb, const float*, int h) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      a[c] = ld4<AAL>(pa + 8 * c + 4 * h);
      b[c] = ld4<true>(pb + 8 * c + 4 * h);
    }
  }
  __device__ inline void mma(floatx16& acc) const {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[c].x, b[c].x, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[c].y, b[c].y, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[c].z, b[c].z, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[c].w, b[c].w, acc, 0, 0, 0);
    }
  }
};

template <>
struct Frag<MM_BF16> {
  bf16x8 a[4], b[4];
  template <bool AAL>
  __device__ inline void load(const __bf16* pa, const __bf16*, const __bf16* pb, const __bf16*, int h) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      a[c] = ld8<AAL>(pa + 16 * c + 8 * h);
      b[c] = ld8<true>(pb + 16 * c + 8 * h);
    }
  }
  __device__ inline void mma(floatx16& acc) const {
#pragma unroll
    for (int c = 0; c < 4; ++c) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[c], b[c], acc, 0, 0, 0);
  }
};

template <>
struct Frag<MM_BF16X3> {
  bf16x8 ah[4], al[4], bh[4], bl[4];
  template <bool AAL>
  __device__ inline void load(const __bf16* pa, const __bf16* pal, const __bf16* pb, const __bf16* pbl, int h) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      ah[c] = ld8<AAL>(pa + 16 * c + 8 * h);
      al[c] = ld8<AAL>(pal + 16 * c + 8 * h);
      bh[c] = ld8<true>(pb + 16 * c + 8 * h);
      bl[c] = ld8<true>(pbl + 16 * c + 8 * h);
    }
  }
  __device__ inline void mma(floatx16& acc) const {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[c], bh[c], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[c], bl[c], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[c], bh[c], acc, 0, 0, 0);
    }
  }
};

// Ring depth per mode (fragment registers: 32 VGPRs per set for FP32 and BF16, 64 for split BF16).
template <int MODE>
constexpr int ring_depth() { return MODE == MM_BF16X3 ? 2 : 4; }

// The part of a reduction of length K (a multiple of the group) that wave `idx` of `parts` owns.
template <int MODE>
__device__ inline void kslice(int64_t K, int parts, int idx, int64_t& k0, int64_t& kl) {
  constexpr int G = kgroup<MODE>();
  const int64_t ng = K / G;
  const int64_t g0 = ng * idx / parts, g1 = ng * (idx + 1) / parts;
  k0 = g0 * G;
  kl = (g1 - g0) * G;
}

// acc += A[r0 + 0..32, k0 .. k0+kl) * B[c0 + 0..32, k0 .. k0+kl)^T for the calling wave.
// kl is a multiple of kgroup<MODE>(); AAL says whether A's rows are 16-byte aligned at k0.
// Fragment set d of the ring is loaded D - 1 groups before its MFMAs; the ring index is a
// compile-time constant after unrolling, so the sets stay in registers.
template <int MODE, bool AAL, typename TA, typename TB>
__device__ inline void wave_mm(floatx16& acc, const Opnd<TA>& A, const Opnd<TB>& B, int64_t r0, int64_t c0, int64_t k0,
                               int64_t kl) {
  if (kl <= 0) return;
  constexpr int G = kgroup<MODE>();
  constexpr int D = ring_depth<MODE>();
  const int lane = threadIdx.x & 63;
  const int rr = lane & 31, h = lane >> 5;
  const TA* a = A.hi + (r0 + rr) * A.ld + k0;
  const TA* al = A.lo + (r0 + rr) * A.ld + k0;
  const TB* b = B.hi + (c0 + rr) * B.ld + k0;
  const TB* bl = B.lo + (c0 + rr) * B.ld + k0;
  Frag<MODE> f[D];
#pragma unroll
  for (int d = 0; d < D - 1; ++d)
    if (d * G < kl) f[d].template load<AAL>(a + d * G, al + d * G, b + d * G, bl + d * G, h);
  for (int64_t g = 0; g < kl; g += D * G) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int64_t kk = g + d * G;
      if (kk < kl) {
        const int64_t kn = kk + (D - 1) * G;
        if (kn < kl) f[(d + D - 1) % D].template load<AAL>(a + kn, al + kn, b + kn, bl + kn, h);
        f[d].mma(acc);
      }
    }
  }
}

// The four waves' 32 x 32 partial accumulators. Register i of lane l holds element
// (8 (i / 4) + 4 (l / 32) + i % 4, l % 32) of the block; rows of 66 words keep the column-wise
// read-back of the epilogues to two-way bank conflicts.
struct Red4 {
  float v[4][16][66];
};

__device__ inline void red_put(Red4& R, const floatx16& acc) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < 16; ++i) R.v[w][i][l] = acc[i];
}

__device__ inline float red_get(const Red4& R, int w, int r, int c) {
  return R.v[w][4 * (r >> 3) + (r & 3)][c + 32 * ((r >> 2) & 1)];
}

__device__ inline float red_sum4(const Red4& R, int r, int c) {
  return (red_get(R, 0, r, c) + red_get(R, 1, r, c)) + (red_get(R, 2, r, c) + red_get(R, 3, r, c));
}

// Four consecutive values at element `off` of a kp x kp matrix: the FP32 copy (if f32) and, with
// SPLIT, the (hi, lo) BF16 pair stored as [hi | lo] blocks of nel elements each (if s16).
template <bool SPLIT>
__device__ inline void put4(float* f32, uint16_t* s16, int64_t nel, int64_t off, const float v[4]) {
  if (f32) *reinterpret_cast<float4*>(f32 + off) = make_float4(v[0], v[1], v[2], v[3]);
  if constexpr (SPLIT) {
    if (s16) {
      bf16x4 hi, lo;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        hi[e] = (__bf16)v[e];
        lo[e] = (__bf16)(v[e] - (float)hi[e]);
      }
      __bf16* s = reinterpret_cast<__bf16*>(s16);
      *reinterpret_cast<bf16x4*>(s + off) = hi;
      *reinterpret_cast<bf16x4*>(s + nel + off) = lo;
    }
  }
}

// Shared memory of the tile epilogue.
struct TileShm {
  float t[32][33];    // the block transposed
  double d[4];        // per-wave ||Out - prev||^2
  float m[4][2][32];  // per-wave partial column sums mu^T Out
};

// Writes block (i0, j0) of a new kp x kp iterate held as v[e] = element (i0 + r, j0 + c4 + e), with
// r = thread / 8 and c4 = 4 (thread % 8): Out row-major, OutT transposed (FP32 and, with SPLIT, BF16
// pairs), frob += ||Out - prev||_F^2 over the block when prev is given, and, when mpart is given,
// mpart[(ti * 2 + s) * kp + j] = sum over the block's rows i of mu_s[i] Out[i][j] (s = A, B).
// Every thread of the block calls it.
template <bool SPLIT>
__device__ inline void finish_tile(TileShm& S, int kp, int ti, int64_t i0, int64_t j0, const float v[4], float* Out,
                                   uint16_t* Out16, float* OutT, uint16_t* OutT16, const float* prev, double* frob,
                                   const double* mu_a, const double* mu_b, float* mpart) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r = threadIdx.x >> 3, c4 = 4 * (threadIdx.x & 7);
  const int64_t nel = (int64_t)kp * kp;
  const int64_t off = (i0 + r) * kp + j0 + c4;
  put4<SPLIT>(Out, Out16, nel, off, v);
#pragma unroll
  for (int e = 0; e < 4; ++e) S.t[c4 + e][r] = v[e];
  if (prev) {
    const float4 p = *reinterpret_cast<const float4*>(prev + off);
    const float pv[4] = {p.x, p.y, p.z, p.w};
    double d = 0.0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const double dd = (double)v[e] - (double)pv[e];
      d += dd * dd;
    }
    for (int o = 32; o > 0; o >>= 1) d += __shfl_xor(d, o);
    if (lane == 0) S.d[w] = d;
  }
  if (mpart) {
    const float ua = (float)mu_a[i0 + r], ub = (float)mu_b[i0 + r];
    float sa[4], sb[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      sa[e] = ua * v[e];
      sb[e] = ub * v[e];
      for (int o = 8; o < 64; o <<= 1) {  // lanes with equal lane % 8 hold the same columns
        sa[e] += __shfl_xor(sa[e], o);
        sb[e] += __shfl_xor(sb[e], o);
      }
    }
    if (lane < 8) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        S.m[w][0][c4 + e] = sa[e];
        S.m[w][1][c4 + e] = sb[e];
      }
    }
  }
  __syncthreads();
  {
    const int c = threadIdx.x >> 3, r4 = 4 * (threadIdx.x & 7);
    const float t[4] = {S.t[c][r4], S.t[c][r4 + 1], S.t[c][r4 + 2], S.t[c][r4 + 3]};
    put4<SPLIT>(OutT, OutT16, nel, (j0 + c) * kp + i0 + r4, t);
  }
  if (prev && threadIdx.x == 0) atomicAdd(frob, (S.d[0] + S.d[1]) + (S.d[2] + S.d[3]));
  if (mpart && threadIdx.x < 64) {
    const int s = threadIdx.x >> 5, c = threadIdx.x & 31;
    mpart[((int64_t)ti * 2 + s) * kp + j0 + c] = (S.m[0][s][c] + S.m[1][s][c]) + (S.m[2][s][c] + S.m[3][s][c]);
  }
}

template <typename T>
__device__ inline void store4(T* p, const float v[4]) {
  if constexpr (sizeof(T) == 2) {
    bf16x4 x;
#pragma unroll
    for (int e = 0; e < 4; ++e) x[e] = (__bf16)v[e];
    *reinterpret_cast<bf16x4*>(p) = x;
  } else {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
  }
}

// ---- P1 --------------------------------------------------------------------------------------------
// Blocks [0, ntb * ntk): a 32-row slice of the minibatch x 32 columns of C_g; the rest: H blocks.
template <int DMODE, int KMODE, typename TY, typename TPH>
__global__ __launch_bounds__(256) void p1_kernel(StgdArgs a, const TY* __restrict__ YA, const TY* __restrict__ YB,
                                                 Opnd<op_t<DMODE>> phiT_d, Opnd<op_t<KMODE>> phiT_k,
                                                 const float* __restrict__ mpart, TPH* __restrict__ cgt, float* __restrict__ h,
                                                 uint16_t* __restrict__ h16, int64_t row0, int32_t nb) {
  __shared__ Red4 red;
  __shared__ float mvec[2][32];
  const int kp = a.kp;
  const int ntk = kp / 32, ntb = a.b / 32;
  const int w = threadIdx.x >> 6;
  int tile = blockIdx.x;
  floatx16 acc = {};
  int64_t k0, kl;
  if (tile < ntb * ntk) {
    const int tb = tile / ntk, tj = tile % ntk;
    const int64_t x0 = (int64_t)tb * 32, j0 = (int64_t)tj * 32;
    // waves 0, 1: P_A over two halves of the k reduction; waves 2, 3: P_B
    const Opnd<TY> Y = opnd<TY>((w < 2 ? YA : YB) + row0 * (int64_t)kp, nullptr, kp);
    kslice<DMODE>(kp, 2, w & 1, k0, kl);
    wave_mm<DMODE, true>(acc, Y, phiT_d, x0, j0, k0, kl);
    if (threadIdx.x < 64) {  // m_A, m_B of the block's columns from the per-row-block partials
      const int s = threadIdx.x >> 5, c = threadIdx.x & 31;
      float m = 0.f;
      for (int ti = 0; ti < ntk; ++ti) m += mpart[((int64_t)ti * 2 + s) * kp + j0 + c];
      mvec[s][c] = m;
    }
    red_put(red, acc);
    __syncthreads();
    // C_g^T rows are columns j of the block: thread -> column c, rows r4 .. r4 + 3
    const int c = threadIdx.x >> 3, r4 = 4 * (threadIdx.x & 7);
    const float mA = mvec[0][c], mB = mvec[1][c];
    float cg[4], cs = 0.f;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int r = r4 + e;
      const float p = red_get(red, 0, r, c) + red_get(red, 1, r, c);
      const float q = red_get(red, 2, r, c) + red_get(red, 3, r, c);
      const bool valid = (x0 + r) < nb;
      const float ab = p * q;
      cs += valid ? ab : 0.f;
      cg[e] = valid ? a.c3 * ab + a.c1 * (p * mB + mA * q) : 0.f;
    }
    store4(cgt + (j0 + c) * (int64_t)a.b + x0 + r4, cg);
    for (int off = 1; off < 8; off <<= 1) cs += __shfl_xor(cs, off);
    if ((threadIdx.x & 7) == 0) {
      a.cs_part[(int64_t)tb * kp + j0 + c] = cs;
      if (tb == 0) {
        a.mab[j0 + c] = mA;
        a.mab[kp + j0 + c] = mB;
      }
    }
  } else {
    tile -= ntb * ntk;
    const int ti = tile / ntk, tj = tile % ntk;
    const int64_t i0 = (int64_t)ti * 32, j0 = (int64_t)tj * 32;
    kslice<KMODE>(kp, 4, w, k0, kl);
    wave_mm<KMODE, true>(acc, phiT_k, phiT_k, i0, j0, k0, kl);  // G = Phi^T Phi
    red_put(red, acc);
    __syncthreads();
    const int r = threadIdx.x >> 3, c4 = 4 * (threadIdx.x & 7);
    float g[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      g[e] = red_sum4(red, r, c4 + e);
      g[e] *= g[e];
    }
    put4<KMODE == MM_BF16X3>(h, h16, (int64_t)kp * kp, (i0 + r) * kp + j0 + c4, g);
  }
}

// ---- P2 --------------------------------------------------------------------------------------------
template <int DMODE, int KMODE, bool YAL, typename TY, typename TPH>
__global__ __launch_bounds__(256) void p2_kernel(StgdArgs a, const TY* __restrict__ YCt, const TPH* __restrict__ cgt,
                                                 Opnd<op_t<KMODE>> phi_k, Opnd<op_t<KMODE>> h_k, uint16_t* __restrict__ phit16,
                                                 float* __restrict__ phitT, uint16_t* __restrict__ phitT16,
                                                 float* __restrict__ rowsum0, int64_t row0, int32_t nb) {
  __shared__ Red4 red;
  __shared__ TileShm S;
  __shared__ float cmu[32];
  const int kp = a.kp, ntk = kp / 32;
  const int w = threadIdx.x >> 6;
  const int ti = blockIdx.x / ntk, tj = blockIdx.x % ntk;
  const int64_t i0 = (int64_t)ti * 32, j0 = (int64_t)tj * 32;
  floatx16 acc = {};
  int64_t k0, kl;
  if (w < 2) {  // Phi H over two halves of k (H symmetric: row j of H is column j)
    kslice<KMODE>(kp, 2, w, k0, kl);
    wave_mm<KMODE, true>(acc, phi_k, h_k, i0, j0, k0, kl);
  } else {      // Y_C^T C_g over two halves of the minibatch tile
    kslice<DMODE>(a.b, 2, w & 1, k0, kl);
    wave_mm<DMODE, YAL>(acc, opnd<TY>(YCt + row0, nullptr, a.n_pad), opnd<TPH>(cgt, nullptr, a.b), i0, j0, k0, kl);
  }
  if (threadIdx.x < 32) {
    const int c = threadIdx.x;
    float s = 0.f;
    for (int tb = 0; tb < a.b / 32; ++tb) s += a.cs_part[(int64_t)tb * kp + j0 + c];
    cmu[c] = a.c0 * a.mab[j0 + c] * a.mab[kp + j0 + c] + a.c1 * s / (float)nb;
  }
  if (blockIdx.x == 0)
    for (int t = threadIdx.x; t < kp; t += blockDim.x) rowsum0[t] = 0.f;
  red_put(red, acc);
  __syncthreads();
  const float eta = a.eta, inv_b = 1.f / (float)nb;
  const int r = threadIdx.x >> 3, c4 = 4 * (threadIdx.x & 7);
  const float muc = (float)a.mu_c[i0 + r];
  const float4 ph = *reinterpret_cast<const float4*>(a.phi + (i0 + r) * kp + j0 + c4);
  const float p0[4] = {ph.x, ph.y, ph.z, ph.w};
  float v[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int c = c4 + e;
    const float phh = red_get(red, 0, r, c) + red_get(red, 1, r, c);
    const float t = red_get(red, 2, r, c) + red_get(red, 3, r, c);
    v[e] = p0[e] - eta * a.theta * phh + eta * (t * inv_b + muc * cmu[c]);
  }
  finish_tile<KMODE == MM_BF16X3>(S, kp, ti, i0, j0, v, a.phit, phit16, phitT, phitT16, nullptr, nullptr, nullptr, nullptr,
                                   nullptr);
}

// ---- Newton-Schulz ---------------------------------------------------------------------------------
// G = X^T X from X^T, and rowsum[i] += sum over the block's columns of |G_ij|.
template <int KMODE>
__global__ __launch_bounds__(256) void ns_gram_kernel(int32_t kp, Opnd<op_t<KMODE>> xT, float* __restrict__ G,
                                                      uint16_t* __restrict__ G16, float* __restrict__ rowsum) {
  __shared__ Red4 red;
  const int ntk = kp / 32;
  const int w = threadIdx.x >> 6;
  const int ti = blockIdx.x / ntk, tj = blockIdx.x % ntk;
  const int64_t i0 = (int64_t)ti * 32, j0 = (int64_t)tj * 32;
  floatx16 acc = {};
  int64_t k0, kl;
  kslice<KMODE>(kp, 4, w, k0, kl);
  wave_mm<KMODE, true>(acc, xT, xT, i0, j0, k0, kl);
  red_put(red, acc);
  __syncthreads();
  const int r = threadIdx.x >> 3, c4 = 4 * (threadIdx.x & 7);
  float g[4], s = 0.f;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    g[e] = red_sum4(red, r, c4 + e);
    s += fabsf(g[e]);
  }
  put4<KMODE == MM_BF16X3>(G, G16, (int64_t)kp * kp, (i0 + r) * kp + j0 + c4, g);
  for (int off = 1; off < 8; off <<= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 7) == 0) atomicAdd(rowsum + i0 + r, s);
}

// Xout = c (1.5 X - 0.5 c^2 X G); the outputs are those of finish_tile.
template <int KMODE>
__global__ __launch_bounds__(256) void ns_update_kernel(int32_t kp, Opnd<op_t<KMODE>> x_k, Opnd<op_t<KMODE>> g_k,
                                                        const float* __restrict__ X, const float* __restrict__ rowsum,
                                                        float* __restrict__ rowsum_next, float* Xout, uint16_t* Xout16,
                                                        float* XoutT, uint16_t* XoutT16, const float* prev, double* frob,
                                                        const double* mu_a, const double* mu_b, float* mpart) {
  __shared__ Red4 red;
  __shared__ TileShm S;
  __shared__ float cscale;
  const int ntk = kp / 32;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int ti = blockIdx.x / ntk, tj = blockIdx.x % ntk;
  const int64_t i0 = (int64_t)ti * 32, j0 = (int64_t)tj * 32;
  floatx16 acc = {};
  int64_t k0, kl;
  kslice<KMODE>(kp, 4, w, k0, kl);
  wave_mm<KMODE, true>(acc, x_k, g_k, i0, j0, k0, kl);  // X G (G symmetric)
  if (w == 3) {  // ||G||_inf from the row sums of |G|
    float mx = 0.f;
    for (int t = lane; t < kp; t += 64) mx = fmaxf(mx, rowsum[t]);
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
    if (lane == 0) cscale = mx > 2.f ? sqrtf(2.f / mx) : 1.f;
  }
  if (blockIdx.x == 0)
    for (int t = threadIdx.x; t < kp; t += blockDim.x) rowsum_next[t] = 0.f;
  red_put(red, acc);
  __syncthreads();
  const float c1 = 1.5f * cscale, c3 = 0.5f * cscale * cscale * cscale;
  const int r = threadIdx.x >> 3, c4 = 4 * (threadIdx.x & 7);
  const float4 xv = *reinterpret_cast<const float4*>(X + (i0 + r) * kp + j0 + c4);
  const float x[4] = {xv.x, xv.y, xv.z, xv.w};
  float v[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) v[e] = c1 * x[e] - c3 * red_sum4(red, r, c4 + e);
  finish_tile<KMODE == MM_BF16X3>(S, kp, ti, i0, j0, v, Xout, Xout16, XoutT, XoutT16, prev, frob, mu_a, mu_b, mpart);
}

// Steps without renormalisation: Phi_next = Phi~, with the outputs of the last NS update. With
// prev == nullptr and Xout == nullptr it is the run set-up: the derived copies of the current Phi.
template <bool SPLIT>
__global__ __launch_bounds__(256) void copy_tile_kernel(int32_t kp, const float* __restrict__ X, float* Xout, uint16_t* Xout16,
                                                        float* XoutT, uint16_t* XoutT16, const float* prev, double* frob,
                                                        const double* mu_a, const double* mu_b, float* mpart) {
  __shared__ TileShm S;
  const int ntk = kp / 32;
  const int ti = blockIdx.x / ntk, tj = blockIdx.x % ntk;
  const int64_t i0 = (int64_t)ti * 32, j0 = (int64_t)tj * 32;
  const int r = threadIdx.x >> 3, c4 = 4 * (threadIdx.x & 7);
  const float4 xv = *reinterpret_cast<const float4*>(X + (i0 + r) * kp + j0 + c4);
  const float v[4] = {xv.x, xv.y, xv.z, xv.w};
  finish_tile<SPLIT>(S, kp, ti, i0, j0, v, Xout, Xout16, XoutT, XoutT16, prev, frob, mu_a, mu_b, mpart);
}

__global__ void zero_f64_kernel(int64_t n, double* p) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.0;
}

__global__ void f32_to_bf16_kernel(int64_t n, const float* __restrict__ in, __bf16* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = (__bf16)in[t];
}

int f32_to_bf16(hipStream_t s, int64_t n, const float* in, bf16_bits* out) {
  if (n <= 0) return MMSB_OK;
  hipLaunchKernelGGL(f32_to_bf16_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, n, in,
                     reinterpret_cast<__bf16*>(out));
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

// ---- host driver -----------------------------------------------------------------------------------
// The operand view of a kp x kp matrix for one mode: FP32 data, or the BF16 (hi, lo) pair.
template <int MODE>
static Opnd<op_t<MODE>> pick(const float* f32, const uint16_t* s16, int64_t nel, int64_t ld) {
  if constexpr (MODE == MM_F32) {
    (void)s16;
    (void)nel;
    return opnd<float>(f32, nullptr, ld);
  } else {
    (void)f32;
    const __bf16* p = reinterpret_cast<const __bf16*>(s16);
    return opnd<__bf16>(p, p + nel, ld);
  }
}

template <int DMODE, int KMODE, typename TY, typename TPH>
static int run_steps(hipStream_t s, const StgdArgs& args, const TY* YA, const TY* YB, const TY* YCt, int64_t step0,
                     int64_t nsteps, int32_t renorm_every, int32_t ns_iters, const StgdBuffers& B, int& cur,
                     double* frob_log, StgdTiming* timing) {
  constexpr bool SPLIT = KMODE == MM_BF16X3;
  constexpr int VEC = 16 / sizeof(TY);  // Y_C^T elements per 16-byte load
  const int kp = args.kp, ntk = kp / 32, ntb = args.b / 32;
  const int64_t nel = (int64_t)kp * kp;
  const int64_t spe = ceil_div(args.n_x, args.bl);
  TPH* cgt = static_cast<TPH*>(B.cgt);
  auto s16 = [&](uint16_t* p) -> uint16_t* { return SPLIT ? p : nullptr; };
  hipLaunchKernelGGL(zero_f64_kernel, dim3((unsigned)ceil_div(nsteps, 256)), dim3(256), 0, s, nsteps, frob_log);
  MMSB_LAUNCH_CHECK();
  hipLaunchKernelGGL(copy_tile_kernel<SPLIT>, dim3(ntk * ntk), dim3(256), 0, s, kp, B.phi[cur], nullptr, s16(B.phi16[cur]),
                     B.phiT, s16(B.phiT16), nullptr, nullptr, args.mu_a, args.mu_b, B.mpart);
  MMSB_LAUNCH_CHECK();
  StgdArgs a = args;
  for (int64_t st = 0; st < nsteps; ++st) {
    const int64_t gstep = step0 + st;
    const int64_t row0 = (gstep % spe) * a.bl;
    const int32_t nb = (int32_t)std::min<int64_t>(a.bl, a.n_x - row0);
    a.phi = B.phi[cur];
    a.phit = B.phit;
    if (timing) MMSB_HIP(hipEventRecord(timing->ev[0], s));
    // the data contraction reads Phi^T as BF16 hi only; the k^3 products read the split pair
    hipLaunchKernelGGL((p1_kernel<DMODE, KMODE, TY, TPH>), dim3(ntb * ntk + ntk * ntk), dim3(256), 0, s, a, YA, YB,
                       pick<DMODE>(B.phiT, B.phiT16, nel, kp), pick<KMODE>(B.phiT, B.phiT16, nel, kp), B.mpart, cgt, B.h,
                       s16(B.h16), row0, nb);
    MMSB_LAUNCH_CHECK();
    if (timing) MMSB_HIP(hipEventRecord(timing->ev[1], s));
    const Opnd<op_t<KMODE>> phi_k = pick<KMODE>(B.phi[cur], B.phi16[cur], nel, kp);
    const Opnd<op_t<KMODE>> h_k = pick<KMODE>(B.h, B.h16, nel, kp);
    if (row0 % VEC == 0)
      hipLaunchKernelGGL((p2_kernel<DMODE, KMODE, true, TY, TPH>), dim3(ntk * ntk), dim3(256), 0, s, a, YCt, cgt, phi_k, h_k,
                         s16(B.phit16), B.phitT, s16(B.phitT16), B.rowsum, row0, nb);
    else
      hipLaunchKernelGGL((p2_kernel<DMODE, KMODE, false, TY, TPH>), dim3(ntk * ntk), dim3(256), 0, s, a, YCt, cgt, phi_k, h_k,
                         s16(B.phit16), B.phitT, s16(B.phitT16), B.rowsum, row0, nb);
    MMSB_LAUNCH_CHECK();
    if (timing) MMSB_HIP(hipEventRecord(timing->ev[2], s));
    float* out = B.phi[cur ^ 1];
    uint16_t* out16 = s16(B.phi16[cur ^ 1]);
    double* frob = frob_log + st;
    const bool renorm = renorm_every > 0 && ((gstep + 1) % renorm_every == 0) && ns_iters > 0;
    if (renorm) {
      // iterates: (phit, phitT) -> (phia, phiaT) -> (phit, phitT) ...; the last writes Phi_next
      float* x = B.phit;
      float* xT = B.phitT;
      uint16_t* x16 = B.phit16;
      uint16_t* xT16 = B.phitT16;
      for (int it = 0; it < ns_iters; ++it) {
        const bool last = it == ns_iters - 1;
        const bool to_a = x == B.phit;
        float* dst = last ? out : (to_a ? B.phia : B.phit);
        uint16_t* dst16 = last ? B.phi16[cur ^ 1] : (to_a ? B.phia16 : B.phit16);
        float* dstT = last ? B.phiT : (to_a ? B.phiaT : B.phitT);
        uint16_t* dstT16 = last ? B.phiT16 : (to_a ? B.phiaT16 : B.phitT16);
        float* rs = B.rowsum + (it & 1) * kp;
        float* rs_next = B.rowsum + ((it + 1) & 1) * kp;
        hipLaunchKernelGGL((ns_gram_kernel<KMODE>), dim3(ntk * ntk), dim3(256), 0, s, kp, pick<KMODE>(xT, xT16, nel, kp), B.gt,
                           s16(B.gt16), rs);
        MMSB_LAUNCH_CHECK();
        hipLaunchKernelGGL((ns_update_kernel<KMODE>), dim3(ntk * ntk), dim3(256), 0, s, kp, pick<KMODE>(x, x16, nel, kp),
                           pick<KMODE>(B.gt, B.gt16, nel, kp), x, rs, rs_next, dst, s16(dst16), dstT, s16(dstT16),
                           last ? a.phi : nullptr, frob, a.mu_a, a.mu_b, last ? B.mpart : nullptr);
        MMSB_LAUNCH_CHECK();
        x = dst;
        xT = dstT;
        x16 = dst16;
        xT16 = dstT16;
      }
    } else {
      hipLaunchKernelGGL(copy_tile_kernel<SPLIT>, dim3(ntk * ntk), dim3(256), 0, s, kp, B.phit, out, out16, B.phiT,
                         s16(B.phiT16), a.phi, frob, a.mu_a, a.mu_b, B.mpart);
      MMSB_LAUNCH_CHECK();
    }
    if (timing) {
      MMSB_HIP(hipEventRecord(timing->ev[3], s));
      MMSB_HIP(hipEventSynchronize(timing->ev[3]));
      for (int ph = 0; ph < 3; ++ph) {
        float ms = 0.f;
        MMSB_HIP(hipEventElapsedTime(&ms, timing->ev[ph], timing->ev[ph + 1]));
        timing->ms[ph] += ms;
      }
    }
    cur ^= 1;
  }
  return MMSB_OK;
}

int stgd_run(hipStream_t s, int32_t precision, const StgdArgs& args, const void* YA, const void* YB, const void* YCt,
             int64_t step0, int64_t nsteps, int32_t renorm_every, int32_t ns_iters, const StgdBuffers& bufs, int& cur,
             double* frob_log, StgdTiming* timing) {
  MMSB_CHECK_ARG(args.kp % 64 == 0 && args.b % 64 == 0 && args.kp >= 64 && args.b >= 64,
                 "stgd: kp and b must be positive multiples of 64 (got %d, %d)", args.kp, args.b);
  MMSB_CHECK_ARG(args.bl >= 1 && args.bl <= args.b, "stgd: logical minibatch %d outside [1, %d]", args.bl, args.b);
  MMSB_CHECK_ARG(args.n_pad % 64 == 0, "stgd: n_pad must be a multiple of 64");
  const int64_t spe = ceil_div(args.n_x, args.bl);
  MMSB_CHECK_ARG(args.n_pad >= (spe - 1) * args.bl + args.b, "stgd: n_pad too small");
  MMSB_CHECK_ARG(nsteps >= 0 && ns_iters >= 0, "stgd: negative step or iteration count");
  if (nsteps == 0) return MMSB_OK;
  if (precision == MMSB_PREC_BF16) {
    MMSB_CHECK_ARG(bufs.phi16[0] && bufs.phi16[1] && bufs.phit16 && bufs.phia16 && bufs.phiT16 && bufs.phitT16 &&
                       bufs.phiaT16 && bufs.gt16 && bufs.h16,
                   "stgd: BF16 mode needs the split operand buffers");
    return run_steps<MM_BF16, MM_BF16X3, __bf16, __bf16>(s, args, static_cast<const __bf16*>(YA), static_cast<const __bf16*>(YB),
                                                         static_cast<const __bf16*>(YCt), step0, nsteps, renorm_every, ns_iters,
                                                         bufs, cur, frob_log, timing);
  }
  MMSB_CHECK_ARG(precision == MMSB_PREC_FP32, "stgd: unknown precision %d", precision);
  return run_steps<MM_F32, MM_F32, float, float>(s, args, static_cast<const float*>(YA), static_cast<const float*>(YB),
                                                 static_cast<const float*>(YCt), step0, nsteps, renorm_every, ns_iters, bufs,
                                                 cur, frob_log, timing);
}

}  // namespace mmsb
