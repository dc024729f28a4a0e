// STGD step kernels, FP32 and BF16 MFMA (the hot loop of GPU_Task, stpm.c:497-790).
//
// One step over a minibatch X_b of nb nodes (layouts in stgd.h):
//   P1  P_A = Y_A[X_b] Phi, P_B = Y_B[X_b] Phi, with m_A = Phi^T mu_A, m_B = Phi^T mu_B:
//       C_g = c3 P_A o P_B + c1 (P_A o m_B + m_A o P_B) (rows beyond nb zero);
//       per-tile column sums of P_A o P_B;  and  H = (Phi^T Phi) o (Phi^T Phi)
//   P2  Phi~ = Phi - eta theta Phi H + eta [Y_C[X_b]^T C_g / nb + mu_C c_mu^T],
//       c_mu = c0 m_A o m_B + c1 sum_x(P_A o P_B) / nb                       (stpm.c:532-691)
//   NS  G = X^T X; X <- c X (3I - c^2 G) / 2 (degree 3), or Q = -(5/4) c^2 G + (3/8) c^4 G^2 and
//       X <- c X (15/8 I + Q) (degree 5), with c = min(1, sqrt(2 / ||G||_inf)), ns_iters times
//       (polar renormalisation in place of the QR of stpm.c:774-786). The polar factor is
//       scale-invariant; c only acts when sigma_max(X)^2 may exceed 2, where the unscaled
//       iterations (convergent for sigma < sqrt 3, resp. sqrt(7/3)) would diverge -- e.g. after a
//       minibatch that holds very high-degree nodes.
// (c3, c1, c0) are the entries of the 4x4 matrices of init_buff (stpm.c:1026-1038).
//
// Work decomposition: every GEMM output is cut into 32 x 32 blocks, one per workgroup of four
// waves. The waves split the reduction dimension (each wave owns its K range) and load their
// operand blocks with whole-cache-line loads, rearranged into the MFMA layout through a small
// per-wave LDS buffer (the staging section below), through a ring of D fragment sets, so that up to
// D groups of loads are in flight ahead of the MFMAs (in P1's data tiles each wave computes both
// Y_A Phi and Y_B Phi over its K range, loading each Phi^T block once for the two). The partial
// accumulators meet in padded LDS tiles that all 256 threads read back for coalesced epilogue
// stores. At k = 500 a phase launches 256-768 workgroups.
// K order inside a fragment: for v_mfma_f32_32x32x2f32 lane half h holds k = 4h + i of a chunk
// of eight k for instruction i (the order of a sum over k is free as long as A and B agree);
// for v_mfma_f32_32x32x16_bf16 lane half h holds the eight contiguous k = 8h .. 8h + 7.
// Precision: FP32 mode multiplies FP32 on v_mfma_f32_32x32x2f32. BF16 mode multiplies the data
// contractions (Y Phi, Y_C^T C_g) as BF16 on v_mfma_f32_32x32x16_bf16 and the k x k x k products
// as split BF16 (x = hi + lo; hi.hi + hi.lo + lo.hi, about 16 significant bits); all accumulate
// in FP32. In BF16 mode every k x k operand is written by the kernel that produces it both in
// FP32 and as its (hi, lo) BF16 pair, so the MFMA loops load ready operands and convert nothing.
// m_A and m_B of the next step are reduced from the tiles of the new Phi where it is written.
// Y_A, Y_B, Y_C^T either hold all of X (row offset row0) or, for checkpointed runs, only the current
// minibatch (row offset 0, Y_C^T with leading dimension b), refilled from CSR before each step by a
// caller-supplied launch (StgdArgs::y_window).
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "common.h"
#include "stgd.h"

namespace cg = cooperative_groups;

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

// Waves per workgroup. Every tile's reduction is split over all of them (kWaves parts of the k or
// minibatch range), which keeps kWaves waves of loads in flight per tile -- with one tile per compute
// unit at k ~ 500, the only source of memory-level parallelism. The epilogues run on the first four
// waves (threads 0..255), which own one element group of four each.
#ifndef MMSB_STGD_WAVES
#define MMSB_STGD_WAVES 4
#endif
constexpr int kWaves = MMSB_STGD_WAVES;
constexpr int kThreads = 64 * kWaves;
static_assert(kWaves == 4 || kWaves == 8, "4 or 8 waves per workgroup");

// ---- operand staging ---------------------------------------------------------------------------
// One reduction group of a wave's MFMA operand is a block of 32 rows x 128 bytes (32 FP32 or 64 BF16
// values per row). Read in the MFMA layout -- lane (r, h) taking row r -- every load instruction
// would touch 32 cache lines and use 32 bytes of each, and with a dozen waves per compute unit
// streaming such blocks the L1 refetches each line for each of its four 32-byte pieces. Instead
// every load instruction reads 8 whole 128-byte row segments (lane 8q + s takes 16-byte segment s of
// row q, plus 8 rows per instruction), and a transpose through the wave's own 4 KiB LDS buffer then
// leaves register c of lane (r, h) with segment 2c + h of row r: k = 8c + 4h .. +3 (FP32) or
// 16c + 8h .. +7 (BF16) of the group, the k order the MFMA sequences below assume for A and B alike
// (the order of a sum over k is free as long as both operands agree). Segment s of row q sits at
// slot s ^ ((q >> 1) & 7) of the buffer row, which keeps both the writes (8 rows x 8 slots per
// instruction) and the reads (32 rows x one segment per lane half) on distinct banks in every group
// of 16 lanes. LDS instructions of one wave execute in order, so the reads see the writes and the next
// block's writes land after the reads; the waves of a workgroup use separate buffers and never wait
// on each other here.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <bool AL, typename T>
__device__ inline u32x4 ld16(const T* p) {
  if constexpr (AL) return *reinterpret_cast<const u32x4*>(p);
  constexpr int E = 16 / sizeof(T);  // rows not 16-byte aligned: element by element
  T e[E];
#pragma unroll
  for (int i = 0; i < E; ++i) e[i] = p[i];
  return __builtin_bit_cast(u32x4, e);
}

// The four coalesced loads of one block; row0 (the block's first row and first k) and ld are
// wave-uniform. Each address is that uniform base plus 8c rows plus one 32-bit per-lane byte offset,
// the form the compiler issues with a scalar base register: one offset VGPR per operand instead of a
// 64-bit address pair per load.
template <bool AL, typename T>
__device__ inline void stage_load(u32x4 v[4], const T* row0, int64_t ld) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t off = ((lane >> 3) * (uint32_t)ld + (lane & 7) * (16 / (uint32_t)sizeof(T))) * (uint32_t)sizeof(T);
  const char* base = reinterpret_cast<const char*>(row0);
#pragma unroll
  for (int c = 0; c < 4; ++c) v[c] = ld16<AL>(reinterpret_cast<const T*>(base + 8 * c * ld * (int64_t)sizeof(T) + off));
}

// The calling wave's index within its workgroup, as a wave-uniform (scalar) value.
__device__ inline int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

__device__ inline u32x4* wave_stage() {
  __shared__ u32x4 buf[kWaves][256];
  return buf[wave_id()];
}

// Rearranges a block loaded by stage_load into the MFMA layout, in place.
__device__ inline void stage_swap(u32x4 v[4]) {
  u32x4* buf = wave_stage();
  const int lane = threadIdx.x & 63;
  const int q = lane >> 3, s = lane & 7, r = lane & 31, h = lane >> 5;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int row = q + 8 * c;
    buf[row * 8 + (s ^ ((row >> 1) & 7))] = v[c];
  }
  asm volatile("" ::: "memory");  // compiler ordering only: the reads follow the writes
#pragma unroll
  for (int c = 0; c < 4; ++c) v[c] = buf[r * 8 + ((2 * c + h) ^ ((r >> 1) & 7))];
  asm volatile("" ::: "memory");  // ... and the next block's writes follow these reads
}

__device__ inline floatx4 as_f32(const u32x4& v) { return __builtin_bit_cast(floatx4, v); }
__device__ inline bf16x8 as_bf16(const u32x4& v) { return __builtin_bit_cast(bf16x8, v); }

// Reduction group of one fragment set: 32 k for FP32, 64 k for the BF16 modes (128 bytes per row).
template <int MODE>
constexpr int kgroup() { return MODE == MM_F32 ? 32 : 64; }

// Fragment set of one reduction group for one wave: load() issues the coalesced loads of its blocks
// (A rows from pa, B rows from pb, both at the group's first k), swap() puts them in the MFMA layout
// once they are needed, mma() accumulates.
template <int MODE>
struct Frag;

template <>
struct Frag<MM_F32> {
  u32x4 a[4], b[4];
  template <bool AAL>
  __device__ inline void load(const float* pa, const float*, int64_t lda, const float* pb, const float*, int64_t ldb) {
    stage_load<AAL>(a, pa, lda);
    stage_load<true>(b, pb, ldb);
  }
  __device__ inline void swap() {
    stage_swap(a);
    stage_swap(b);
  }
  __device__ inline void mma(floatx16& acc) const {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const floatx4 x = as_f32(a[c]), y = as_f32(b[c]);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(x[0], y[0], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(x[1], y[1], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(x[2], y[2], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(x[3], y[3], acc, 0, 0, 0);
    }
  }
};

template <>
struct Frag<MM_BF16> {
  u32x4 a[4], b[4];
  template <bool AAL>
  __device__ inline void load(const __bf16* pa, const __bf16*, int64_t lda, const __bf16* pb, const __bf16*,
                              int64_t ldb) {
    stage_load<AAL>(a, pa, lda);
    stage_load<true>(b, pb, ldb);
  }
  __device__ inline void swap() {
    stage_swap(a);
    stage_swap(b);
  }
  __device__ inline void mma(floatx16& acc) const {
#pragma unroll
    for (int c = 0; c < 4; ++c) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16(a[c]), as_bf16(b[c]), acc, 0, 0, 0);
  }
};

template <>
struct Frag<MM_BF16X3> {
  u32x4 ah[4], al[4], bh[4], bl[4];
  template <bool AAL>
  __device__ inline void load(const __bf16* pa, const __bf16* pal, int64_t lda, const __bf16* pb, const __bf16* pbl,
                              int64_t ldb) {
    stage_load<AAL>(ah, pa, lda);
    stage_load<AAL>(al, pal, lda);
    stage_load<true>(bh, pb, ldb);
    stage_load<true>(bl, pbl, ldb);
  }
  __device__ inline void swap() {
    stage_swap(ah);
    stage_swap(al);
    stage_swap(bh);
    stage_swap(bl);
  }
  __device__ inline void mma(floatx16& acc) const {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const bf16x8 xh = as_bf16(ah[c]), xl = as_bf16(al[c]), yh = as_bf16(bh[c]), yl = as_bf16(bl[c]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, yh, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, yl, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xl, yh, acc, 0, 0, 0);
    }
  }
};

// Ring depth per mode (fragment registers: 32 VGPRs per set for FP32 and BF16, 64 for split BF16).
// Two sets keep 8 (FP32, BF16) or 16 (split BF16) loads per wave in flight ahead of the MFMAs.
#ifndef MMSB_RING_F32
#define MMSB_RING_F32 2
#endif
#ifndef MMSB_RING_BF16
#define MMSB_RING_BF16 2
#endif
#ifndef MMSB_RING_BF16X3
#define MMSB_RING_BF16X3 2
#endif
static_assert(MMSB_RING_F32 >= 2 && MMSB_RING_BF16 >= 2 && MMSB_RING_BF16X3 >= 2, "ring depth must be at least 2");
template <int MODE>
constexpr int ring_depth() {
  return MODE == MM_BF16X3 ? MMSB_RING_BF16X3 : (MODE == MM_BF16 ? MMSB_RING_BF16 : MMSB_RING_F32);
}

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
  const TA* a = A.hi + r0 * A.ld + k0;
  const TA* al = A.lo + r0 * A.ld + k0;
  const TB* b = B.hi + c0 * B.ld + k0;
  const TB* bl = B.lo + c0 * B.ld + k0;
  Frag<MODE> f[D];
#pragma unroll
  for (int d = 0; d < D - 1; ++d)
    if (d * G < kl) f[d].template load<AAL>(a + d * G, al + d * G, A.ld, b + d * G, bl + d * G, B.ld);
  for (int64_t g = 0; g < kl; g += D * G) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int64_t kk = g + d * G;
      if (kk < kl) {
        const int64_t kn = kk + (D - 1) * G;
        if (kn < kl) f[(d + D - 1) % D].template load<AAL>(a + kn, al + kn, A.ld, b + kn, bl + kn, B.ld);
        f[d].swap();
        f[d].mma(acc);
      }
    }
  }
}

// Two products that share their column operand, acc_a += A1 B^T and acc_b += A2 B^T, for one wave
// over k in [k0, k0 + kl) (FP32 or BF16 modes): the fragment set holds A1, A2 and B, so B is loaded
// once for both.
template <int MODE>
struct FragPair {
  using T = op_t<MODE>;
  u32x4 a1[4], a2[4], b[4];
  __device__ inline void load(const T* p1, const T* p2, int64_t lda, const T* pb, int64_t ldb) {
    stage_load<true>(a1, p1, lda);
    stage_load<true>(a2, p2, lda);
    stage_load<true>(b, pb, ldb);
  }
  __device__ inline void swap() {
    stage_swap(a1);
    stage_swap(a2);
    stage_swap(b);
  }
  __device__ inline void mma(floatx16& acc_a, floatx16& acc_b) const {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if constexpr (MODE == MM_F32) {
        const floatx4 x1 = as_f32(a1[c]), x2 = as_f32(a2[c]), y = as_f32(b[c]);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          acc_a = __builtin_amdgcn_mfma_f32_32x32x2f32(x1[e], y[e], acc_a, 0, 0, 0);
          acc_b = __builtin_amdgcn_mfma_f32_32x32x2f32(x2[e], y[e], acc_b, 0, 0, 0);
        }
      } else {
        acc_a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16(a1[c]), as_bf16(b[c]), acc_a, 0, 0, 0);
        acc_b = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16(a2[c]), as_bf16(b[c]), acc_b, 0, 0, 0);
      }
    }
  }
};

template <int MODE>
__device__ inline void wave_mm_pair(floatx16& acc_a, floatx16& acc_b, const op_t<MODE>* A1, const op_t<MODE>* A2, int64_t lda,
                                    const op_t<MODE>* B, int64_t ldb, int64_t r0, int64_t c0, int64_t k0, int64_t kl) {
  if (kl <= 0) return;
  constexpr int G = kgroup<MODE>();
  constexpr int D = 2;
  const op_t<MODE>* a1 = A1 + r0 * lda + k0;
  const op_t<MODE>* a2 = A2 + r0 * lda + k0;
  const op_t<MODE>* b = B + c0 * ldb + k0;
  FragPair<MODE> f[D];
  f[0].load(a1, a2, lda, b, ldb);
  for (int64_t g = 0; g < kl; g += D * G) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int64_t kk = g + d * G;
      if (kk < kl) {
        const int64_t kn = kk + (D - 1) * G;
        if (kn < kl) f[(d + D - 1) % D].load(a1 + kn, a2 + kn, lda, b + kn, ldb);
        f[d].swap();
        f[d].mma(acc_a, acc_b);
      }
    }
  }
}

__device__ inline bool epi_thread() { return threadIdx.x < 256; }

// Four 32 x 32 partial accumulators. Register i of lane l holds element
// (8 (i / 4) + 4 (l / 32) + i % 4, l % 32) of the block; rows of 66 words keep the column-wise
// read-back of the epilogues to two-way bank conflicts.
struct Red4 {
  float v[4][16][66];
};

__device__ inline void red_slot_put(Red4& R, int slot, const floatx16& acc) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < 16; ++i) R.v[slot][i][l] = acc[i];
}

// Leaves the wave partials in the four slots of R: with eight waves, wave w + 4 adds its accumulator
// to wave w's first (fixed order), so slot w holds acc_w + acc_{w+4}. The caller synchronises before
// reading R. Every thread of the block calls it.
__device__ inline void red_put(Red4& R, floatx16 acc) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if constexpr (kWaves == 8) {
    if (w >= 4) red_slot_put(R, w - 4, acc);
    __syncthreads();
    if (w < 4) {
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[i] += R.v[w][i][l];
    }
  }
  if (w < 4) red_slot_put(R, w, acc);
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

// Inputs of finish_tile that do not depend on the tile's products, loaded before the MFMA loop so
// that their latency overlaps the operand loads: the previous iterate's block values (for the frob
// log) and mu_A, mu_B of the thread's row (for the m partial sums).
struct TilePre {
  float pv[4];
  float ua, ub;
};

__device__ inline void tile_pre(TilePre& t, int kp, int64_t i0, int64_t j0, const float* prev, const double* mu_a,
                                const double* mu_b) {
  if (!epi_thread()) return;
  const int r = threadIdx.x >> 3, c4 = 4 * (threadIdx.x & 7);
  if (prev) {
    const float4 p = *reinterpret_cast<const float4*>(prev + (i0 + r) * kp + j0 + c4);
    t.pv[0] = p.x;
    t.pv[1] = p.y;
    t.pv[2] = p.z;
    t.pv[3] = p.w;
  }
  if (mu_a) {
    t.ua = (float)mu_a[i0 + r];
    t.ub = (float)mu_b[i0 + r];
  }
}

// Shared memory of the tile epilogue.
struct TileShm {
  float t[32][33];    // the block transposed
  double d[4];        // per-wave ||Out - prev||^2
  float m[4][2][32];  // per-wave partial column sums mu^T Out
};

// Writes block (i0, j0) of a new kp x kp iterate held as v[e] = element (i0 + r, j0 + c4 + e), with
// r = thread / 8 and c4 = 4 (thread % 8): Out row-major (FP32 and, with SPLIT, BF16 pairs), OutT
// transposed (FP32 without SPLIT, BF16 pairs with it: the transposes are only ever MFMA operands, and
// a later FP32 run rebuilds its FP32 transpose in SETUP), *fslot = ||Out - prev||_F^2 over the block when fslot is given (prev in pre.pv; the step's
// log entry is the sum of its tiles' slots, see frob_sum), and, when
// mpart is given, mpart[(ti * 2 + s) * kp + j] = sum over the block's rows i of mu_s[i] Out[i][j]
// (s = A, B; mu in pre.ua, pre.ub). Every thread of the block calls it.
template <bool SPLIT>
__device__ inline void finish_tile(TileShm& S, int kp, int ti, int64_t i0, int64_t j0, const float v[4], float* Out,
                                   uint16_t* Out16, float* OutT, uint16_t* OutT16, const TilePre& pre, double* fslot,
                                   float* mpart) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r = threadIdx.x >> 3, c4 = 4 * (threadIdx.x & 7);
  const int64_t nel = (int64_t)kp * kp;
  const int64_t off = (i0 + r) * kp + j0 + c4;
  const bool act = epi_thread();  // waves 4.. hold no elements; they only meet the barrier
  if (act) {
  put4<SPLIT>(Out, Out16, nel, off, v);
#pragma unroll
  for (int e = 0; e < 4; ++e) S.t[c4 + e][r] = v[e];
  if (fslot) {
    double d = 0.0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const double dd = (double)v[e] - (double)pre.pv[e];
      d += dd * dd;
    }
    for (int o = 32; o > 0; o >>= 1) d += __shfl_xor(d, o);
    if (lane == 0) S.d[w] = d;
  }
  if (mpart) {
    float sa[4], sb[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      sa[e] = pre.ua * v[e];
      sb[e] = pre.ub * v[e];
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
  }
  __syncthreads();
  if (act) {
    const int c = threadIdx.x >> 3, r4 = 4 * (threadIdx.x & 7);
    const float t[4] = {S.t[c][r4], S.t[c][r4 + 1], S.t[c][r4 + 2], S.t[c][r4 + 3]};
    put4<SPLIT>(SPLIT ? nullptr : OutT, OutT16, nel, (j0 + c) * kp + i0 + r4, t);
  }
  if (fslot && threadIdx.x == 0) *fslot = (S.d[0] + S.d[1]) + (S.d[2] + S.d[3]);
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

// ---- the phases of one step, as per-tile device functions ----------------------------------------
// A step is the phase sequence P1, P2, then ns_iters x (GRAM, UPDATE) -- (GRAM, QPOLY, UPDATE) for
// degree-5 iterations -- or COPY when the step does not renormalise; SETUP precedes the first step of
// a run and FROB follows its last. In the reference's QR mode a renormalising step is P1, P2, the
// host's QR launches and QRFIN. The kernel that finishes a step leaves one partial of the step's log
// entry per tile; the next step's P1 (or FROB) sums them in a fixed order. The same tile functions
// serve the per-phase launches of one instance (stgd_phase_kernel, blockIdx.x = tile) and the
// persistent launch over many instances (stgd_batch_kernel, tiles from a work queue).
enum StgdPhase { PH_SETUP = 0, PH_P1 = 1, PH_P2 = 2, PH_GRAM = 3, PH_UPDATE = 4, PH_COPY = 5, PH_QPOLY = 6, PH_QRFIN = 7,
                 PH_FROB = 8 };

// Everything the phases of one step of one instance read or write.
struct StepCtx {
  StgdArgs a;          // a.phi = B.phi[cur], a.phit = B.phit
  StgdBuffers B;
  const void* YA;
  const void* YB;
  const void* YCt;
  int32_t cur;         // B.phi[cur] holds the step's input Phi
  int32_t ns_iters;
  int32_t ns_degree;   // 3 or 5
  int64_t row0;        // first X row of the minibatch
  int64_t yrow0;       // its first row in the Y buffers (row0, or 0 when they hold only the minibatch)
  int32_t nb;          // rows in the minibatch
  int64_t st;          // step index within the run
  double* frob_prev;   // the previous step's log entry within the run (null for the run's first), formed by
                       // this step's P1 from the previous step's tile partials
  double* frob;        // this step's ||Phi_{t+1} - Phi_t||_F^2 (written by FROB after a run's last step)
  double* frob_zero;   // SETUP: the run's frob entries, cleared by tile 0
  int64_t nfrob;
};

union StgdSmem {
  struct {
    Red4 red[2];  // data tiles: the four waves' partials of P_A (red[0]) and P_B (red[1])
    float mvec[2][32];
  } p1;
  struct {
    Red4 red;
    TileShm S;
    float cmu[32];
    float cscale;
  } p2;
  TileShm S;
};

// Tiles of a phase. The symmetric products -- H = (Phi^T Phi)^o2 in P1, G = X^T X in GRAM -- are
// computed on the upper triangle of 32 x 32 blocks only, each off-diagonal block written twice.
__host__ __device__ inline int phase_tiles(int phase, int kp, int b) {
  const int ntk = kp / 32, tri = ntk * (ntk + 1) / 2;
  switch (phase) {
    case PH_P1: return (b / 32) * ntk + tri;
    case PH_GRAM: return tri;
    case PH_FROB: return 1;
    default: return ntk * ntk;
  }
}

// Block (ti, tj), ti <= tj, of upper-triangle tile t (row-major over the triangle).
__device__ inline void tri_tile(int t, int ntk, int& ti, int& tj) {
  int i = 0;
  while (t >= ntk - i) {
    t -= ntk - i;
    ++i;
  }
  ti = i;
  tj = i + t;
}

// The log entry ||Phi_{t+1} - Phi_t||_F^2 of a finished step from its tiles' partials fpart[0 .. n):
// lane-strided sums, then a butterfly -- a fixed order, so that every wave that forms the entry gets
// the same bits. Called by whole waves.
__device__ inline double frob_sum(const double* fpart, int n) {
  const int lane = threadIdx.x & 63;
  double s = 0.0;
  for (int t = lane; t < n; t += 64) s += fpart[t];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  return s;
}

// Newton-Schulz iteration `it` reads (x, xT) and writes (dst, dstT): phit -> phia -> phit -> ...;
// the last iteration writes phi[cur ^ 1] and phiT.
struct NsBufs {
  float *x, *xT, *dst, *dstT;
  uint16_t *x16, *xT16, *dst16, *dstT16;
};

__host__ __device__ inline NsBufs ns_bufs(const StgdBuffers& B, int cur, int it, int ns_iters) {
  NsBufs n;
  const bool from_t = (it & 1) == 0;
  const bool last = it == ns_iters - 1;
  n.x = from_t ? B.phit : B.phia;
  n.xT = from_t ? B.phitT : B.phiaT;
  n.x16 = from_t ? B.phit16 : B.phia16;
  n.xT16 = from_t ? B.phitT16 : B.phiaT16;
  n.dst = last ? B.phi[cur ^ 1] : (from_t ? B.phia : B.phit);
  n.dstT = last ? B.phiT : (from_t ? B.phiaT : B.phitT);
  n.dst16 = last ? B.phi16[cur ^ 1] : (from_t ? B.phia16 : B.phit16);
  n.dstT16 = last ? B.phiT16 : (from_t ? B.phiaT16 : B.phitT16);
  return n;
}

// The operand view of a kp x kp matrix for one mode: FP32 data, or the BF16 (hi, lo) pair.
template <int MODE>
__host__ __device__ inline Opnd<op_t<MODE>> pick(const float* f32, const uint16_t* s16, int64_t nel, int64_t ld) {
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

template <bool SPLIT>
__device__ inline uint16_t* sp(uint16_t* p) {
  return SPLIT ? p : nullptr;
}

// Sums of strided partials, one per output column: out[o] = sum_{t < nt} src[t * stride + off(o)]
// for 64 outputs (P1: m_A, m_B of 32 columns) or 32 (P2: column sums). The loads of all threads
// are issued by part_load before the MFMA loop and reduced by part_reduce after it, so their
// latency overlaps the operand loads instead of forming a chain.
constexpr int kPartRegs = 4;
struct PartSum {
  float v[kPartRegs];
  int extra;  // partials beyond kPartRegs * lanes-per-output, summed late (large k or b)
};

// P1: tile (ti rows of Phi) partials of m_s[j0 + c]: thread -> output o = t / 4 (s = o / 32,
// c = o % 32), lane q = t % 4 takes ti = q, q + 4, ...
__device__ inline void mpart_load(PartSum& ps, const float* mpart, int kp, int ntk, int64_t j0) {
  const int o = threadIdx.x >> 2, q = threadIdx.x & 3;
  const int s = o >> 5, c = o & 31;
#pragma unroll
  for (int u = 0; u < kPartRegs; ++u) {
    const int ti = q + 4 * u;
    ps.v[u] = ti < ntk ? mpart[((int64_t)ti * 2 + s) * kp + j0 + c] : 0.f;
  }
  ps.extra = ntk > 4 * kPartRegs;
}

__device__ inline float mpart_reduce(const PartSum& ps, const float* mpart, int kp, int ntk, int64_t j0) {
  const int o = threadIdx.x >> 2, q = threadIdx.x & 3;
  const int s = o >> 5, c = o & 31;
  float m = (ps.v[0] + ps.v[1]) + (ps.v[2] + ps.v[3]);
  if (ps.extra)
    for (int ti = q + 4 * kPartRegs; ti < ntk; ti += 4) m += mpart[((int64_t)ti * 2 + s) * kp + j0 + c];
  m += __shfl_xor(m, 1);
  m += __shfl_xor(m, 2);
  return m;  // valid in every lane of the output's group of four
}

// P2: column sums of the b/32 row-tile partials cs_part[tb][j0 + c]: thread -> c = t / 8,
// lane q = t % 8 takes tb = q, q + 8, ...
__device__ inline void cspart_load(PartSum& ps, const float* cs_part, int kp, int ntb, int64_t j0) {
  const int c = threadIdx.x >> 3, q = threadIdx.x & 7;
#pragma unroll
  for (int u = 0; u < kPartRegs; ++u) {
    const int tb = q + 8 * u;
    ps.v[u] = tb < ntb ? cs_part[(int64_t)tb * kp + j0 + c] : 0.f;
  }
  ps.extra = ntb > 8 * kPartRegs;
}

__device__ inline float cspart_reduce(const PartSum& ps, const float* cs_part, int kp, int ntb, int64_t j0) {
  const int c = threadIdx.x >> 3, q = threadIdx.x & 7;
  float m = (ps.v[0] + ps.v[1]) + (ps.v[2] + ps.v[3]);
  if (ps.extra)
    for (int tb = q + 8 * kPartRegs; tb < ntb; tb += 8) m += cs_part[(int64_t)tb * kp + j0 + c];
  m += __shfl_xor(m, 1);
  m += __shfl_xor(m, 2);
  m += __shfl_xor(m, 4);
  return m;
}

// P1 tiles [0, ntk (ntk + 1) / 2): upper-triangle H tiles; the rest: a 32-row slice of the minibatch x
// 32 columns of C_g.
template <int DMODE, int KMODE, typename TY, typename TPH>
__device__ __forceinline__ void p1_tile(const StepCtx& P, int tile, StgdSmem& sm) {
  constexpr bool SPLIT = KMODE == MM_BF16X3;
  const StgdArgs& a = P.a;
  const int kp = a.kp;
  const int ntk = kp / 32;
  const int64_t nel = (int64_t)kp * kp;
  const int w = wave_id();
  const int tri = ntk * (ntk + 1) / 2;
  floatx16 acc = {};
  int64_t k0, kl;
  if (tile >= tri) {  // data tiles follow the H tiles, which then start in the first wave
    tile -= tri;
    const int tb = tile / ntk, tj = tile % ntk;
    const int64_t x0 = (int64_t)tb * 32, j0 = (int64_t)tj * 32;
    PartSum ps;
    if (epi_thread()) mpart_load(ps, P.B.mpart, kp, ntk, j0);  // m_A, m_B of the block's columns
    // every wave: P_A and P_B over its part of the k reduction, sharing the Phi^T fragments
    floatx16 acc_b = {};
    kslice<DMODE>(kp, kWaves, w, k0, kl);
    wave_mm_pair<DMODE>(acc, acc_b, static_cast<const TY*>(P.YA) + P.yrow0 * (int64_t)kp,
                        static_cast<const TY*>(P.YB) + P.yrow0 * (int64_t)kp, kp, pick<DMODE>(P.B.phiT, P.B.phiT16, nel, kp).hi,
                        kp, x0, j0, k0, kl);
    if (epi_thread()) {
      const float m = mpart_reduce(ps, P.B.mpart, kp, ntk, j0);
      const int o = threadIdx.x >> 2;
      if ((threadIdx.x & 3) == 0) sm.p1.mvec[o >> 5][o & 31] = m;
    }
    red_put(sm.p1.red[0], acc);
    red_put(sm.p1.red[1], acc_b);
    __syncthreads();
    if (!epi_thread()) return;
    // C_g^T rows are columns j of the block: thread -> column c, rows r4 .. r4 + 3
    const int c = threadIdx.x >> 3, r4 = 4 * (threadIdx.x & 7);
    const float mA = sm.p1.mvec[0][c], mB = sm.p1.mvec[1][c];
    float cg[4], cs = 0.f;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int r = r4 + e;
      const float p = red_sum4(sm.p1.red[0], r, c);
      const float q = red_sum4(sm.p1.red[1], r, c);
      const bool valid = (x0 + r) < P.nb;
      const float ab = p * q;
      cs += valid ? ab : 0.f;
      cg[e] = valid ? a.c3 * ab + a.c1 * (p * mB + mA * q) : 0.f;
    }
    store4(static_cast<TPH*>(P.B.cgt) + (j0 + c) * (int64_t)a.b + x0 + r4, cg);
    for (int off = 1; off < 8; off <<= 1) cs += __shfl_xor(cs, off);
    if ((threadIdx.x & 7) == 0) {
      a.cs_part[(int64_t)tb * kp + j0 + c] = cs;
      if (tb == 0) {
        a.mab[j0 + c] = mA;
        a.mab[kp + j0 + c] = mB;
      }
    }
  } else {
    int ti, tj;
    tri_tile(tile, ntk, ti, tj);
    const int64_t i0 = (int64_t)ti * 32, j0 = (int64_t)tj * 32;
    const Opnd<op_t<KMODE>> phiT_k = pick<KMODE>(P.B.phiT, P.B.phiT16, nel, kp);
    kslice<KMODE>(kp, kWaves, w, k0, kl);
    wave_mm<KMODE, true>(acc, phiT_k, phiT_k, i0, j0, k0, kl);  // G = Phi^T Phi
    Red4& red = sm.p1.red[0];
    red_put(red, acc);
    __syncthreads();
    if (!epi_thread()) return;
    const int r = threadIdx.x >> 3, c4 = 4 * (threadIdx.x & 7);
    float g[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      g[e] = red_sum4(red, r, c4 + e);
      g[e] *= g[e];
    }
    put4<SPLIT>(P.B.h, sp<SPLIT>(P.B.h16), nel, (i0 + r) * kp + j0 + c4, g);
    if (ti != tj) {  // the mirror block (j, i): thread -> row c of it, columns r4 .. r4 + 3
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        g[e] = red_sum4(red, c4 + e, r);
        g[e] *= g[e];
      }
      put4<SPLIT>(P.B.h, sp<SPLIT>(P.B.h16), nel, (j0 + r) * kp + i0 + c4, g);
    }
  }
}

template <int DMODE, int KMODE, typename TY, typename TPH>
__device__ __forceinline__ void p2_tile(const StepCtx& P, int tile, StgdSmem& sm) {
  constexpr bool SPLIT = KMODE == MM_BF16X3;
  constexpr int VEC = 16 / sizeof(TY);  // Y_C^T elements per 16-byte load
  const StgdArgs& a = P.a;
  const int kp = a.kp, ntk = kp / 32;
  const int64_t nel = (int64_t)kp * kp;
  const int w = wave_id();
  const int ti = tile / ntk, tj = tile % ntk;
  const int64_t i0 = (int64_t)ti * 32, j0 = (int64_t)tj * 32;
  Red4& red = sm.p2.red;
  floatx16 acc = {};
  int64_t k0, kl;
  PartSum ps;  // column sums of P_A o P_B and m_A, m_B for c_mu, gathered during the MFMA loop
  const bool epi = epi_thread();
  const int r = threadIdx.x >> 3, c4 = 4 * (threadIdx.x & 7);
  float mab_a = 0.f, mab_b = 0.f, muc = 0.f;
  float4 ph = make_float4(0.f, 0.f, 0.f, 0.f);
  if (epi) {  // epilogue inputs, loaded ahead of the MFMA loop
    cspart_load(ps, a.cs_part, kp, a.b / 32, j0);
    mab_a = a.mab[j0 + (threadIdx.x >> 3)];
    mab_b = a.mab[kp + j0 + (threadIdx.x >> 3)];
    muc = (float)a.mu_c[i0 + r];
    ph = *reinterpret_cast<const float4*>(a.phi + (i0 + r) * kp + j0 + c4);
  }
  // waves with (w & 2) == 0 take Phi H, the others Y_C^T C_g, each over kWaves / 2 parts of its range;
  // after red_put, slots 0, 1 hold Phi H and slots 2, 3 Y_C^T C_g
  const int part = (w & 1) + 2 * (w >> 2);
  if ((w & 2) == 0) {  // Phi H (H symmetric: row j of H is column j)
    kslice<KMODE>(kp, kWaves / 2, part, k0, kl);
    wave_mm<KMODE, true>(acc, pick<KMODE>(a.phi, P.B.phi16[P.cur], nel, kp), pick<KMODE>(P.B.h, P.B.h16, nel, kp), i0, j0,
                         k0, kl);
  } else {             // Y_C^T C_g over parts of the minibatch tile
    kslice<DMODE>(a.b, kWaves / 2, part, k0, kl);
    const Opnd<TY> yc = opnd<TY>(static_cast<const TY*>(P.YCt) + P.yrow0, nullptr, a.n_pad);
    const Opnd<TPH> cg = opnd<TPH>(static_cast<const TPH*>(P.B.cgt), nullptr, a.b);
    if (P.yrow0 % VEC == 0)
      wave_mm<DMODE, true>(acc, yc, cg, i0, j0, k0, kl);
    else
      wave_mm<DMODE, false>(acc, yc, cg, i0, j0, k0, kl);
  }
  if (epi) {
    const int c = threadIdx.x >> 3;
    const float ssum = cspart_reduce(ps, a.cs_part, kp, a.b / 32, j0);
    if ((threadIdx.x & 7) == 0) sm.p2.cmu[c] = a.c0 * mab_a * mab_b + a.c1 * ssum / (float)P.nb;
  }
  if (tile == 0)
    for (int t = threadIdx.x; t < kp; t += blockDim.x) P.B.rowsum[t] = 0ull;
  red_put(red, acc);
  __syncthreads();
  const float eta = a.eta, inv_b = 1.f / (float)P.nb;
  const float p0[4] = {ph.x, ph.y, ph.z, ph.w};
  float v[4] = {0.f, 0.f, 0.f, 0.f};
  if (epi) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int c = c4 + e;
      const float phh = red_get(red, 0, r, c) + red_get(red, 1, r, c);
      const float t = red_get(red, 2, r, c) + red_get(red, 3, r, c);
      v[e] = p0[e] - eta * a.theta * phh + eta * (t * inv_b + muc * sm.p2.cmu[c]);
    }
  }
  TilePre pre{};  // the reference's log: ||Phi~ - Phi||^2 before the renormalisation (frob_pre)
#pragma unroll
  for (int e = 0; e < 4; ++e) pre.pv[e] = p0[e];
  finish_tile<SPLIT>(sm.p2.S, kp, ti, i0, j0, v, a.phit, sp<SPLIT>(P.B.phit16), P.B.phitT, sp<SPLIT>(P.B.phitT16),
                     pre, a.frob_pre ? P.B.fpart + tile : nullptr, nullptr);
}

// The Newton-Schulz scale guard c = min(1, sqrt(2 / ||G||_inf)) from the row sums of |G|, computed
// by wave 3 of a block: guard_load issues the loads before the block's MFMA loop, guard_scale
// reduces them after it and leaves c in *cscale (shared memory; read after the next barrier).
// The row sums are accumulated across tiles as fixed-point integers (units of 2^-32): integer
// addition is associative, so the sums -- and the guard they decide -- do not depend on the order in
// which the tiles' atomics land, and Phi is bit-reproducible from run to run. Partial sums are
// clamped at 2^30, far beyond the point where the guard acts.
__device__ inline unsigned long long fix32(float s) {
  return (unsigned long long)(long long)llrintf(fminf(s, 1073741824.f) * 4294967296.f);
}

__device__ inline float unfix32(unsigned long long v) { return (float)((double)v * 2.3283064365386963e-10); }

__device__ inline void guard_load(float rs[8], const unsigned long long* rowsum, int kp) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int u = 0; u < 8; ++u) rs[u] = lane + 64 * u < kp ? unfix32(rowsum[lane + 64 * u]) : 0.f;
}

__device__ inline void guard_scale(const float rs[8], const unsigned long long* rowsum, int kp, float* cscale) {
  const int lane = threadIdx.x & 63;
  float mx = 0.f;
#pragma unroll
  for (int u = 0; u < 8; ++u) mx = fmaxf(mx, rs[u]);
  for (int t = lane + 512; t < kp; t += 64) mx = fmaxf(mx, unfix32(rowsum[t]));
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
  if (lane == 0) *cscale = mx > 2.f ? sqrtf(2.f / mx) : 1.f;
}

// G = X^T X of Newton-Schulz iteration `it` on upper-triangle tile `tile` (block and mirror), and
// rowsum[i] += sum over the block's columns of |G_ij| for every row i the block and its mirror cover.
template <int KMODE>
__device__ __forceinline__ void gram_tile(const StepCtx& P, int it, int tile, StgdSmem& sm) {
  constexpr bool SPLIT = KMODE == MM_BF16X3;
  const int kp = P.a.kp, ntk = kp / 32;
  const int64_t nel = (int64_t)kp * kp;
  const int w = wave_id();
  int ti, tj;
  tri_tile(tile, ntk, ti, tj);
  const int64_t i0 = (int64_t)ti * 32, j0 = (int64_t)tj * 32;
  const NsBufs nb = ns_bufs(P.B, P.cur, it, P.ns_iters);
  unsigned long long* rowsum = P.B.rowsum + (it & 1) * kp;
  Red4& red = sm.p2.red;
  floatx16 acc = {};
  int64_t k0, kl;
  const Opnd<op_t<KMODE>> xT = pick<KMODE>(nb.xT, nb.xT16, nel, kp);
  kslice<KMODE>(kp, kWaves, w, k0, kl);
  wave_mm<KMODE, true>(acc, xT, xT, i0, j0, k0, kl);
  red_put(red, acc);
  __syncthreads();
  if (!epi_thread()) return;
  const int r = threadIdx.x >> 3, c4 = 4 * (threadIdx.x & 7);
  float g[4], s = 0.f;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    g[e] = red_sum4(red, r, c4 + e);
    s += fabsf(g[e]);
  }
  put4<SPLIT>(P.B.gt, sp<SPLIT>(P.B.gt16), nel, (i0 + r) * kp + j0 + c4, g);
  for (int off = 1; off < 8; off <<= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 7) == 0) atomicAdd(rowsum + i0 + r, fix32(s));
  if (ti != tj) {  // the mirror block (j, i): thread -> row r of it, columns c4 .. c4 + 3
    float t = 0.f;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      g[e] = red_sum4(red, c4 + e, r);
      t += fabsf(g[e]);
    }
    put4<SPLIT>(P.B.gt, sp<SPLIT>(P.B.gt16), nel, (j0 + r) * kp + i0 + c4, g);
    for (int off = 1; off < 8; off <<= 1) t += __shfl_xor(t, off);
    if ((threadIdx.x & 7) == 0) atomicAdd(rowsum + j0 + r, fix32(t));
  }
}

// Degree-5 iteration, second stage: Q = -(5/4) c^2 G + (3/8) c^4 G^2 (symmetric, as G), so that the
// update is Xout = c X (15/8 I + Q).
template <int KMODE>
__device__ __forceinline__ void qpoly_tile(const StepCtx& P, int it, int tile, StgdSmem& sm) {
  constexpr bool SPLIT = KMODE == MM_BF16X3;
  const int kp = P.a.kp, ntk = kp / 32;
  const int64_t nel = (int64_t)kp * kp;
  const int w = wave_id();
  const int ti = tile / ntk, tj = tile % ntk;
  const int64_t i0 = (int64_t)ti * 32, j0 = (int64_t)tj * 32;
  const unsigned long long* rowsum = P.B.rowsum + (it & 1) * kp;
  Red4& red = sm.p2.red;
  floatx16 acc = {};
  int64_t k0, kl;
  kslice<KMODE>(kp, kWaves, w, k0, kl);
  float rs[8];
  if (w == 3) guard_load(rs, rowsum, kp);
  const int r = threadIdx.x >> 3, c4 = 4 * (threadIdx.x & 7);
  float4 gv = make_float4(0.f, 0.f, 0.f, 0.f);
  if (epi_thread()) gv = *reinterpret_cast<const float4*>(P.B.gt + (i0 + r) * kp + j0 + c4);
  const Opnd<op_t<KMODE>> g = pick<KMODE>(P.B.gt, P.B.gt16, nel, kp);
  wave_mm<KMODE, true>(acc, g, g, i0, j0, k0, kl);  // (G G)_ij = G[i,:] . G[j,:] (G symmetric)
  if (w == 3) guard_scale(rs, rowsum, kp, &sm.p2.cscale);
  red_put(red, acc);
  __syncthreads();
  if (!epi_thread()) return;
  const float c2 = sm.p2.cscale * sm.p2.cscale;
  const float a1 = -1.25f * c2, a2 = 0.375f * c2 * c2;
  const float gg[4] = {gv.x, gv.y, gv.z, gv.w};
  float v[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) v[e] = a1 * gg[e] + a2 * red_sum4(red, r, c4 + e);
  put4<SPLIT>(P.B.q, sp<SPLIT>(P.B.q16), nel, (i0 + r) * kp + j0 + c4, v);
}

// Degree 3: Xout = c (1.5 X - 0.5 c^2 X G); degree 5: Xout = c (1.875 X + X Q). The outputs are those
// of finish_tile (the last iteration also logs ||Phi_next - Phi||^2 and the partial sums of m_A, m_B).
template <int KMODE>
__device__ __forceinline__ void update_tile(const StepCtx& P, int it, int tile, StgdSmem& sm) {
  constexpr bool SPLIT = KMODE == MM_BF16X3;
  const int kp = P.a.kp, ntk = kp / 32;
  const int64_t nel = (int64_t)kp * kp;
  const int w = wave_id();
  const int ti = tile / ntk, tj = tile % ntk;
  const int64_t i0 = (int64_t)ti * 32, j0 = (int64_t)tj * 32;
  const NsBufs nb = ns_bufs(P.B, P.cur, it, P.ns_iters);
  const bool last = it == P.ns_iters - 1;
  const unsigned long long* rowsum = P.B.rowsum + (it & 1) * kp;
  unsigned long long* rowsum_next = P.B.rowsum + ((it + 1) & 1) * kp;
  Red4& red = sm.p2.red;
  floatx16 acc = {};
  int64_t k0, kl;
  kslice<KMODE>(kp, kWaves, w, k0, kl);
  float rs[8];
  if (w == 3) guard_load(rs, rowsum, kp);
  const bool quintic = P.ns_degree == 5;
  const int r = threadIdx.x >> 3, c4 = 4 * (threadIdx.x & 7);
  float4 xv = make_float4(0.f, 0.f, 0.f, 0.f);
  if (epi_thread()) xv = *reinterpret_cast<const float4*>(nb.x + (i0 + r) * kp + j0 + c4);
  TilePre pre;
  tile_pre(pre, kp, i0, j0, last ? P.a.phi : nullptr, last ? P.a.mu_a : nullptr, P.a.mu_b);
  wave_mm<KMODE, true>(acc, pick<KMODE>(nb.x, nb.x16, nel, kp),
                       quintic ? pick<KMODE>(P.B.q, P.B.q16, nel, kp) : pick<KMODE>(P.B.gt, P.B.gt16, nel, kp), i0, j0,
                       k0, kl);  // X G or X Q (both symmetric)
  if (w == 3) guard_scale(rs, rowsum, kp, &sm.p2.cscale);
  if (tile == 0)
    for (int t = threadIdx.x; t < kp; t += blockDim.x) rowsum_next[t] = 0ull;
  red_put(red, acc);
  __syncthreads();
  const float cs = sm.p2.cscale;
  const float c1 = (quintic ? 1.875f : 1.5f) * cs, c3 = quintic ? cs : -0.5f * cs * cs * cs;
  const float x[4] = {xv.x, xv.y, xv.z, xv.w};
  float v[4] = {0.f, 0.f, 0.f, 0.f};
  if (epi_thread())
#pragma unroll
    for (int e = 0; e < 4; ++e) v[e] = c1 * x[e] + c3 * red_sum4(red, r, c4 + e);
  finish_tile<SPLIT>(sm.p2.S, kp, ti, i0, j0, v, nb.dst, sp<SPLIT>(nb.dst16), nb.dstT, sp<SPLIT>(nb.dstT16), pre,
                     last ? P.B.fpart + tile : nullptr, last ? P.B.mpart : nullptr);
}

// COPY: Phi_next = Phi~ with the outputs of a last update; QRFIN: Phi_next = the QR factor the host
// launches left in phia (rows and columns from k on set to zero), without a log entry (the step logged
// before its QR). SETUP: the derived copies of the current Phi (split pairs, Phi^T, the m partial
// sums), and tile 0 clears the run's frob entries.
template <bool SPLIT>
__device__ __forceinline__ void copy_tile(const StepCtx& P, bool setup, int tile, StgdSmem& sm, bool qrfin = false) {
  const int kp = P.a.kp, ntk = kp / 32;
  const int ti = tile / ntk, tj = tile % ntk;
  const int64_t i0 = (int64_t)ti * 32, j0 = (int64_t)tj * 32;
  const int r = threadIdx.x >> 3, c4 = 4 * (threadIdx.x & 7);
  const float* X = setup ? P.B.phi[P.cur] : (qrfin ? P.B.phia : P.B.phit);
  float4 xv = make_float4(0.f, 0.f, 0.f, 0.f);
  if (epi_thread()) xv = *reinterpret_cast<const float4*>(X + (i0 + r) * kp + j0 + c4);
  float v[4] = {xv.x, xv.y, xv.z, xv.w};
  if (qrfin)
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (i0 + r >= P.a.k || j0 + c4 + e >= P.a.k) v[e] = 0.f;
  if (setup && tile == 0)
    for (int64_t t = threadIdx.x; t < P.nfrob; t += blockDim.x) P.frob_zero[t] = 0.0;
  TilePre pre;
  tile_pre(pre, kp, i0, j0, setup ? nullptr : P.a.phi, P.a.mu_a, P.a.mu_b);
  if (setup)
    finish_tile<SPLIT>(sm.S, kp, ti, i0, j0, v, nullptr, sp<SPLIT>(P.B.phi16[P.cur]), P.B.phiT, sp<SPLIT>(P.B.phiT16),
                       pre, nullptr, P.B.mpart);
  else
    finish_tile<SPLIT>(sm.S, kp, ti, i0, j0, v, P.B.phi[P.cur ^ 1], sp<SPLIT>(P.B.phi16[P.cur ^ 1]), P.B.phiT,
                       sp<SPLIT>(P.B.phiT16), pre, (qrfin || P.a.frob_pre) ? nullptr : P.B.fpart + tile, P.B.mpart);
}

// FROB: the log entry of a run's last step from its tiles' partials (one wave).
__device__ __forceinline__ void frob_final(const StepCtx& P) {
  if (threadIdx.x >= 64) return;
  const int ntk = P.a.kp / 32;
  const double f = frob_sum(P.B.fpart, ntk * ntk);
  if (threadIdx.x == 0) *P.frob = f;
}

// P1's tile 0 also forms the previous step's log entry (wave 0, after its tile work).
__device__ __forceinline__ void p1_log(const StepCtx& P) {
  if (!P.frob_prev || threadIdx.x >= 64) return;
  const int ntk = P.a.kp / 32;
  const double f = frob_sum(P.B.fpart, ntk * ntk);
  if (threadIdx.x == 0) *P.frob_prev = f;
}

template <int DMODE, int KMODE, typename TY, typename TPH>
__device__ __forceinline__ void run_tile(const StepCtx& P, int phase, int it, int tile, StgdSmem& sm) {
  constexpr bool SPLIT = KMODE == MM_BF16X3;
  switch (phase) {
    case PH_SETUP: copy_tile<SPLIT>(P, true, tile, sm); break;
    case PH_P1:
      p1_tile<DMODE, KMODE, TY, TPH>(P, tile, sm);
      if (tile == 0) p1_log(P);
      break;
    case PH_FROB: frob_final(P); break;
    case PH_P2: p2_tile<DMODE, KMODE, TY, TPH>(P, tile, sm); break;
    case PH_GRAM: gram_tile<KMODE>(P, it, tile, sm); break;
    case PH_UPDATE: update_tile<KMODE>(P, it, tile, sm); break;
    case PH_QPOLY: qpoly_tile<KMODE>(P, it, tile, sm); break;
    case PH_QRFIN: copy_tile<SPLIT>(P, false, tile, sm, true); break;
    default: copy_tile<SPLIT>(P, false, tile, sm); break;
  }
}

// The end test of the reference's STGD loop after step t (sgd_iter, 0-based) with log value f
// (stpm.c:747-765): converged, or past the cap.
__device__ inline bool stop_test(const StgdArgs& a, int64_t t, double f) {
  return (a.iter_min > 0 && f < a.thresh && t > a.iter_min) || (a.iter_max > 0 && t > a.iter_max);
}

// One phase of one step, one kernel per phase (each with its own register allocation):
// blockIdx.x = tile. (An XCD-aware order that gave each XCD a compact group of tiles cut the HBM
// fetches of P1 from 24 to 9 MiB per launch, but the step took 2.6 % longer: the waves of a group
// then request the same lines at the same moment.)
// With the end test on (P.a.stop), every phase of step g = base + st first checks whether the loop
// has ended; P1 of each step but a run's first forms the previous step's log entry in every wave
// (its kernels have finished), workgroup 0 stores it, and every workgroup applies the test to it and
// marks the end -- each reaches the same decision from the same bits, so the whole launch returns
// together. P1 returns early only when an earlier step ended the loop (stop_at < g): stop_at == g
// was set by a sibling workgroup, and workgroup 0 must still store the entry. The skipped kernels
// write nothing, and the host reads stop_at after the run.
template <int DMODE, int KMODE, typename TY, typename TPH, int PHASE>
__global__ __launch_bounds__(kThreads) void stgd_phase_kernel(StepCtx P, int32_t it) {
  __shared__ StgdSmem sm;
  constexpr bool SPLIT = KMODE == MM_BF16X3;
  if constexpr (PHASE != PH_SETUP) {
    if (P.a.stop) {
      const int64_t g = P.a.stop->base + P.st;
      const int64_t sa = P.a.stop->stop_at;
      if (sa < g || (PHASE != PH_P1 && sa == g)) return;  // an earlier step ended the loop (checked
                                                           // first: the log of a skipped step is never written)
      if constexpr (PHASE == PH_P1) {
        if (P.frob_prev) {
          const int ntk = P.a.kp / 32;
          const double f = frob_sum(P.B.fpart, ntk * ntk);
          if (blockIdx.x == 0 && threadIdx.x == 0) *P.frob_prev = f;
          if (stop_test(P.a, g - 1, f)) {
            if (threadIdx.x == 0) P.a.stop->stop_at = g;
            return;
          }
        }
      }
    }
  }
  if constexpr (PHASE == PH_SETUP) copy_tile<SPLIT>(P, true, blockIdx.x, sm);
  else if constexpr (PHASE == PH_P1) {
    p1_tile<DMODE, KMODE, TY, TPH>(P, blockIdx.x, sm);
    if (!P.a.stop && blockIdx.x == 0) p1_log(P);
  } else if constexpr (PHASE == PH_FROB) frob_final(P);
  else if constexpr (PHASE == PH_P2) p2_tile<DMODE, KMODE, TY, TPH>(P, blockIdx.x, sm);
  else if constexpr (PHASE == PH_GRAM) gram_tile<KMODE>(P, it, blockIdx.x, sm);
  else if constexpr (PHASE == PH_UPDATE) update_tile<KMODE>(P, it, blockIdx.x, sm);
  else if constexpr (PHASE == PH_QPOLY) qpoly_tile<KMODE>(P, it, blockIdx.x, sm);
  else if constexpr (PHASE == PH_QRFIN) copy_tile<SPLIT>(P, false, blockIdx.x, sm, true);
  else copy_tile<SPLIT>(P, false, blockIdx.x, sm);
}

// ---- many instances in one persistent launch -------------------------------------------------------
// The host lays the phases of every instance's steps out slot by slot: slot s holds the s-th phase
// segment of every instance that has one, so each segment's inputs were produced in an earlier
// slot. The cooperative grid walks the slots; in each, the workgroups share the slot's tiles
// (tile t of the slot goes to workgroup t mod grid), and a grid barrier separates consecutive
// slots. Every workgroup executes every barrier, and no workgroup waits on another except there.
struct BatchSeg {
  int32_t inst;
  int32_t phase;
  int32_t it;
  int32_t ntiles;
  int64_t step;     // step index within the launch
  int64_t first;    // index of the segment's first tile within its slot
  int64_t ctx;      // its step's context in the plan's context array
};

struct BatchInst {
  StgdArgs a;
  StgdBuffers B;
  const void* YA;
  const void* YB;
  const void* YCt;
  int32_t cur0;       // phi[cur0] holds Phi at the start of the launch
  int32_t ns_iters;
  int32_t ns_degree;
  int64_t step0;      // global index of the launch's first step
  int64_t spe;        // steps per epoch
  double* frob_log;   // [steps of the launch]
  int64_t nfrob;
};

__host__ __device__ inline void make_step(const BatchInst& I, int64_t st, StepCtx& P) {
  P.a = I.a;
  P.B = I.B;
  P.YA = I.YA;
  P.YB = I.YB;
  P.YCt = I.YCt;
  P.cur = I.cur0 ^ (int)(st & 1);
  P.ns_iters = I.ns_iters;
  P.ns_degree = I.ns_degree;
  const int64_t g = I.step0 + st;
  P.row0 = (g % I.spe) * I.a.bl;
  P.yrow0 = P.row0;
  P.nb = (int32_t)min((int64_t)I.a.bl, I.a.n_x - P.row0);
  P.st = st;
  P.frob_prev = st > 0 ? I.frob_log + st - 1 : nullptr;
  P.a.phi = I.B.phi[P.cur];
  P.a.phit = I.B.phit;
  P.frob = I.frob_log + st;
  P.frob_zero = I.frob_log;
  P.nfrob = I.nfrob;
}

// The segment of a slot that holds slot tile t (the last with first <= t). Every lane runs the same
// search on the same addresses, so it stays in scalar registers.
__device__ inline int32_t find_seg(const BatchSeg* segs, int32_t s0, int32_t s1, int64_t t) {
  int32_t lo = s0, hi = s1 - 1;
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (segs[mid].first <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// The step contexts are built on the host with the plan, one per (member, step), and read here through
// a uniform reference into global memory, as the per-phase kernels read theirs from the kernel
// arguments: in scalar registers. (A context held in LDS is per-lane data to the compiler, which costs
// the scalar operand bases and doubles the register footprint; one built in registers goes to the
// stack, as it is indexed at run time. The tile functions are forced inline for the same reason: an
// out-of-line call passes the context as a per-lane pointer.)
template <int DMODE, int KMODE, typename TY, typename TPH>
__global__ __launch_bounds__(kThreads) void stgd_batch_kernel(const StepCtx* __restrict__ ctxs, const BatchSeg* __restrict__ segs,
                                                         const int32_t* slot_seg, const int64_t* slot_tiles,
                                                         int32_t nslots) {
  cg::grid_group grid = cg::this_grid();
  __shared__ StgdSmem sm;
  for (int32_t sl = 0; sl < nslots; ++sl) {
    const int32_t s0 = slot_seg[sl], s1 = slot_seg[sl + 1];
    const int64_t ntiles = slot_tiles[sl];
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
      const BatchSeg sg = segs[find_seg(segs, s0, s1, t)];
      run_tile<DMODE, KMODE, TY, TPH>(ctxs[sg.ctx], sg.phase, sg.it, (int)(t - sg.first), sm);
      __syncthreads();
    }
    grid.sync();
  }
}

// One phase slot of the plan as its own launch (blockIdx.x = tile of the slot): the grouped
// alternative to the persistent kernel, where the kernel boundary orders consecutive slots.
template <int DMODE, int KMODE, typename TY, typename TPH>
__global__ __launch_bounds__(kThreads) void stgd_slot_kernel(const StepCtx* __restrict__ ctxs, const BatchSeg* __restrict__ segs, int32_t s0, int32_t s1) {
  __shared__ StgdSmem sm;
  const int64_t t = blockIdx.x;
  const BatchSeg sg = segs[find_seg(segs, s0, s1, t)];
  run_tile<DMODE, KMODE, TY, TPH>(ctxs[sg.ctx], sg.phase, sg.it, (int)(t - sg.first), sm);
}

// A slot whose segments are all of one phase, launched with that phase's code only: the generic slot
// kernel holds every phase and gets the register budget of the largest (one wave per SIMD), this one
// that of its phase, as the per-phase kernels of single instances.
template <int DMODE, int KMODE, typename TY, typename TPH, int PHASE>
__global__ __launch_bounds__(kThreads) void stgd_slot_phase_kernel(const StepCtx* __restrict__ ctxs, const BatchSeg* __restrict__ segs, int32_t s0,
                                                                  int32_t s1) {
  __shared__ StgdSmem sm;
  constexpr bool SPLIT = KMODE == MM_BF16X3;
  const int64_t t = blockIdx.x;
  const BatchSeg sg = segs[find_seg(segs, s0, s1, t)];
  const StepCtx& P = ctxs[sg.ctx];
  const int tile = (int)(t - sg.first);
  if constexpr (PHASE == PH_SETUP) copy_tile<SPLIT>(P, true, tile, sm);
  else if constexpr (PHASE == PH_P1) {
    p1_tile<DMODE, KMODE, TY, TPH>(P, tile, sm);
    if (tile == 0) p1_log(P);
  } else if constexpr (PHASE == PH_P2) p2_tile<DMODE, KMODE, TY, TPH>(P, tile, sm);
  else if constexpr (PHASE == PH_GRAM) gram_tile<KMODE>(P, sg.it, tile, sm);
  else if constexpr (PHASE == PH_UPDATE) update_tile<KMODE>(P, sg.it, tile, sm);
  else if constexpr (PHASE == PH_QPOLY) qpoly_tile<KMODE>(P, sg.it, tile, sm);
  else if constexpr (PHASE == PH_FROB) frob_final(P);
  else copy_tile<SPLIT>(P, false, tile, sm);
}

__global__ void set_i64_kernel(int64_t* p, int64_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

int stgd_set_i64(hipStream_t s, int64_t* p, int64_t v) {
  hipLaunchKernelGGL(set_i64_kernel, dim3(1), dim3(64), 0, s, p, v);
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
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
static bool step_renorms(int64_t gstep, int32_t renorm_every, int32_t ns_iters) {
  return renorm_every > 0 && ((gstep + 1) % renorm_every == 0) && ns_iters > 0;
}

template <int DMODE, int KMODE, typename TY, typename TPH>
static int run_steps(hipStream_t s, const StgdArgs& args, const void* YA, const void* YB, const void* YCt, int64_t step0,
                     int64_t nsteps, int32_t renorm_every, int32_t ns_iters, int32_t ns_degree, const StgdBuffers& B,
                     int& cur, double* frob_log, StgdTiming* timing, const StgdFill& fill, const StgdRenorm& renorm) {
  const int kp = args.kp;
  const int64_t spe = ceil_div(args.n_x, args.bl);
  StepCtx P{};
  P.a = args;
  P.B = B;
  P.YA = YA;
  P.YB = YB;
  P.YCt = YCt;
  P.ns_iters = ns_iters;
  P.ns_degree = ns_degree;
  auto launch = [&](int phase, int it) -> int {
    const dim3 grid(phase_tiles(phase, kp, args.b)), block(kThreads);
    switch (phase) {
      case PH_SETUP: hipLaunchKernelGGL((stgd_phase_kernel<DMODE, KMODE, TY, TPH, PH_SETUP>), grid, block, 0, s, P, it); break;
      case PH_P1: hipLaunchKernelGGL((stgd_phase_kernel<DMODE, KMODE, TY, TPH, PH_P1>), grid, block, 0, s, P, it); break;
      case PH_P2: hipLaunchKernelGGL((stgd_phase_kernel<DMODE, KMODE, TY, TPH, PH_P2>), grid, block, 0, s, P, it); break;
      case PH_GRAM: hipLaunchKernelGGL((stgd_phase_kernel<DMODE, KMODE, TY, TPH, PH_GRAM>), grid, block, 0, s, P, it); break;
      case PH_UPDATE: hipLaunchKernelGGL((stgd_phase_kernel<DMODE, KMODE, TY, TPH, PH_UPDATE>), grid, block, 0, s, P, it); break;
      case PH_QPOLY: hipLaunchKernelGGL((stgd_phase_kernel<DMODE, KMODE, TY, TPH, PH_QPOLY>), grid, block, 0, s, P, it); break;
      case PH_QRFIN: hipLaunchKernelGGL((stgd_phase_kernel<DMODE, KMODE, TY, TPH, PH_QRFIN>), grid, block, 0, s, P, it); break;
      case PH_FROB: hipLaunchKernelGGL((stgd_phase_kernel<DMODE, KMODE, TY, TPH, PH_FROB>), grid, block, 0, s, P, it); break;
      default: hipLaunchKernelGGL((stgd_phase_kernel<DMODE, KMODE, TY, TPH, PH_COPY>), grid, block, 0, s, P, it); break;
    }
    MMSB_LAUNCH_CHECK();
    return MMSB_OK;
  };
  P.cur = cur;
  P.a.phi = B.phi[cur];
  P.a.phit = B.phit;
  P.frob_zero = frob_log;
  P.nfrob = nsteps;
  MMSB_TRY(launch(PH_SETUP, 0));
  for (int64_t st = 0; st < nsteps; ++st) {
    const int64_t gstep = step0 + st;
    P.cur = cur;
    P.row0 = (gstep % spe) * args.bl;
    P.nb = (int32_t)std::min<int64_t>(args.bl, args.n_x - P.row0);
    P.yrow0 = args.y_window ? 0 : P.row0;
    P.a.phi = B.phi[cur];
    P.st = st;
    P.frob_prev = st > 0 ? frob_log + st - 1 : nullptr;
    P.frob = frob_log + st;
    if (timing) MMSB_HIP(hipEventRecord(timing->ev[0], s));
    if (args.y_window) MMSB_TRY(fill(P.row0, P.nb));
    MMSB_TRY(launch(PH_P1, 0));
    if (timing) MMSB_HIP(hipEventRecord(timing->ev[1], s));
    MMSB_TRY(launch(PH_P2, 0));
    if (timing) MMSB_HIP(hipEventRecord(timing->ev[2], s));
    if (renorm && renorm_every > 0 && (gstep + 1) % renorm_every == 0) {
      MMSB_TRY(renorm(B.phit, B.phia));
      MMSB_TRY(launch(PH_QRFIN, 0));
    } else if (!renorm && step_renorms(gstep, renorm_every, ns_iters)) {
      for (int it = 0; it < ns_iters; ++it) {
        MMSB_TRY(launch(PH_GRAM, it));
        if (ns_degree == 5) MMSB_TRY(launch(PH_QPOLY, it));
        MMSB_TRY(launch(PH_UPDATE, it));
      }
    } else {
      MMSB_TRY(launch(PH_COPY, 0));
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
  P.st = nsteps - 1;  // the last step's log entry
  P.frob = frob_log + nsteps - 1;
  MMSB_TRY(launch(PH_FROB, 0));
  return MMSB_OK;
}

static int check_args(const StgdArgs& args, int32_t ns_iters, int32_t ns_degree) {
  MMSB_CHECK_ARG(args.kp % 64 == 0 && args.b % 64 == 0 && args.kp >= 64 && args.b >= 64,
                 "stgd: kp and b must be positive multiples of 64 (got %d, %d)", args.kp, args.b);
  MMSB_CHECK_ARG(args.bl >= 1 && args.bl <= args.b, "stgd: logical minibatch %d outside [1, %d]", args.bl, args.b);
  MMSB_CHECK_ARG(args.n_pad % 64 == 0, "stgd: n_pad must be a multiple of 64");
  const int64_t spe = ceil_div(args.n_x, args.bl);
  MMSB_CHECK_ARG(args.n_pad >= (args.y_window ? (int64_t)args.b : (spe - 1) * args.bl + args.b), "stgd: n_pad too small");
  MMSB_CHECK_ARG(ns_iters >= 0, "stgd: negative iteration count");
  MMSB_CHECK_ARG(ns_degree == 3 || ns_degree == 5, "stgd: Newton-Schulz degree must be 3 or 5 (got %d)", ns_degree);
  return MMSB_OK;
}

static int check_split(const StgdBuffers& bufs) {
  MMSB_CHECK_ARG(bufs.phi16[0] && bufs.phi16[1] && bufs.phit16 && bufs.phia16 && bufs.phiT16 && bufs.phitT16 &&
                     bufs.phiaT16 && bufs.gt16 && bufs.h16 && bufs.q16,
                 "stgd: BF16 mode needs the split operand buffers");
  return MMSB_OK;
}

int stgd_run(hipStream_t s, int32_t precision, const StgdArgs& args, const void* YA, const void* YB, const void* YCt,
             int64_t step0, int64_t nsteps, int32_t renorm_every, int32_t ns_iters, int32_t ns_degree,
             const StgdBuffers& bufs, int& cur, double* frob_log, StgdTiming* timing, const StgdFill& fill,
             const StgdRenorm& renorm) {
  MMSB_TRY(check_args(args, ns_iters, ns_degree));
  MMSB_CHECK_ARG(nsteps >= 0, "stgd: negative step count");
  MMSB_CHECK_ARG(!args.y_window || fill, "stgd: a minibatch window needs a fill callback");
  if (nsteps == 0) return MMSB_OK;
  if (precision == MMSB_PREC_BF16) {
    MMSB_TRY(check_split(bufs));
    return run_steps<MM_BF16, MM_BF16X3, __bf16, __bf16>(s, args, YA, YB, YCt, step0, nsteps, renorm_every, ns_iters,
                                                         ns_degree, bufs, cur, frob_log, timing, fill, renorm);
  }
  MMSB_CHECK_ARG(precision == MMSB_PREC_FP32, "stgd: unknown precision %d", precision);
  return run_steps<MM_F32, MM_F32, float, float>(s, args, YA, YB, YCt, step0, nsteps, renorm_every, ns_iters, ns_degree,
                                                 bufs, cur, frob_log, timing, fill, renorm);
}

// ---- batch ------------------------------------------------------------------------------------------
struct StgdBatch {
  int32_t precision = MMSB_PREC_FP32;
  int32_t ninst = 0;
  int32_t nslots = 0;
  int32_t grid = 0;
  int32_t grid_max = 0;
  StepCtx* d_ctx = nullptr;        // [contexts]: one per (member, step)
  BatchSeg* d_segs = nullptr;      // [segments]
  int32_t* d_slot_seg = nullptr;   // [nslots + 1]: first segment of each slot
  int64_t* d_slot_tiles = nullptr; // [nslots]
  std::vector<int32_t> h_slot_seg;  // host copies for the per-slot launches
  std::vector<int64_t> h_slot_tiles;
  std::vector<int32_t> h_slot_phase;  // the phase of every segment of the slot, or -1 when they differ
};

template <int DMODE, int KMODE, typename TY, typename TPH>
static const void* batch_fn() {
  return reinterpret_cast<const void*>(stgd_batch_kernel<DMODE, KMODE, TY, TPH>);
}

static const void* batch_kernel(int32_t precision) {
  return precision == MMSB_PREC_BF16 ? batch_fn<MM_BF16, MM_BF16X3, __bf16, __bf16>() : batch_fn<MM_F32, MM_F32, float, float>();
}

int stgd_batch_plan(StgdBatch*& out, int32_t precision, const std::vector<StgdBatchMember>& members, int32_t ns_iters,
                    int32_t ns_degree, int32_t renorm_every) {
  out = nullptr;
  MMSB_CHECK_ARG(!members.empty(), "stgd batch: no instances");
  MMSB_CHECK_ARG(precision == MMSB_PREC_FP32 || precision == MMSB_PREC_BF16, "stgd batch: unknown precision %d", precision);
  std::vector<BatchInst> insts(members.size());
  std::vector<std::vector<BatchSeg>> per(members.size());
  std::vector<StepCtx> ctxs;                  // one per (member, step)
  std::vector<int64_t> ctx0(members.size());  // each member's first
  for (size_t i = 0; i < members.size(); ++i) {
    const StgdBatchMember& m = members[i];
    MMSB_TRY(check_args(m.args, ns_iters, ns_degree));
    MMSB_CHECK_ARG(!m.args.y_window, "stgd batch: instance %zu recomputes its minibatches (not supported in a batch)", i);
    MMSB_CHECK_ARG(!m.args.stop && !m.args.frob_pre, "stgd batch: instance %zu uses the per-step end test or QR mode", i);
    if (precision == MMSB_PREC_BF16) MMSB_TRY(check_split(m.bufs));
    MMSB_CHECK_ARG(m.nsteps >= 1, "stgd batch: instance %zu has no steps", i);
    BatchInst& I = insts[i];
    I.a = m.args;
    I.B = m.bufs;
    I.YA = m.YA;
    I.YB = m.YB;
    I.YCt = m.YCt;
    I.cur0 = m.cur;
    I.ns_iters = ns_iters;
    I.ns_degree = ns_degree;
    I.step0 = m.step0;
    I.spe = ceil_div(m.args.n_x, m.args.bl);
    I.frob_log = m.frob_log;
    I.nfrob = m.nsteps;
    ctx0[i] = (int64_t)ctxs.size();
    for (int64_t st = 0; st < m.nsteps; ++st) {
      StepCtx P{};
      make_step(I, st, P);
      ctxs.push_back(P);
    }
    auto add = [&](int phase, int it, int64_t st) {
      BatchSeg sg{};
      sg.inst = (int32_t)i;
      sg.phase = phase;
      sg.it = it;
      sg.ntiles = phase_tiles(phase, m.args.kp, m.args.b);
      sg.step = st;
      sg.ctx = ctx0[i] + st;
      per[i].push_back(sg);
    };
    add(PH_SETUP, 0, 0);
    for (int64_t st = 0; st < m.nsteps; ++st) {
      add(PH_P1, 0, st);
      add(PH_P2, 0, st);
      if (step_renorms(m.step0 + st, renorm_every, ns_iters)) {
        for (int it = 0; it < ns_iters; ++it) {
          add(PH_GRAM, it, st);
          if (ns_degree == 5) add(PH_QPOLY, it, st);
          add(PH_UPDATE, it, st);
        }
      } else {
        add(PH_COPY, 0, st);
      }
    }
  }
  std::vector<BatchSeg> segs;
  std::vector<int32_t> slot_seg;
  std::vector<int64_t> slot_tiles;
  std::vector<int32_t> slot_phase;
  // every member's log entry of its last step in one final slot: nothing writes a member's tile
  // partials after its last step, and the other slots keep one phase each
  for (size_t i = 0; i < members.size(); ++i) {
    BatchSeg sg{};
    sg.inst = (int32_t)i;
    sg.phase = PH_FROB;
    sg.ntiles = 1;
    sg.step = members[i].nsteps - 1;
    sg.ctx = ctx0[i] + sg.step;
    per[i].push_back(sg);
  }
  size_t maxlen = 0;
  for (auto& v : per) maxlen = std::max(maxlen, v.size() - 1);
  for (size_t sl = 0; sl <= maxlen; ++sl) {
    slot_seg.push_back((int32_t)segs.size());
    int64_t t = 0;
    int32_t phase = -2;
    for (size_t i = 0; i < per.size(); ++i) {
      const size_t last = per[i].size() - 1;  // the FROB segment
      const size_t at = sl == maxlen ? last : sl;
      if (sl < maxlen && sl >= last) continue;
      BatchSeg sg = per[i][at];
      sg.first = t;
      t += sg.ntiles;
      segs.push_back(sg);
      phase = phase == -2 ? sg.phase : (phase == sg.phase ? phase : -1);
    }
    slot_tiles.push_back(t);
    slot_phase.push_back(phase < 0 ? -1 : phase);
  }
  slot_seg.push_back((int32_t)segs.size());
  MMSB_CHECK_ARG(segs.size() < (size_t)INT32_MAX, "stgd batch: too many segments");
  auto* b = new StgdBatch();
  b->precision = precision;
  b->ninst = (int32_t)insts.size();
  b->nslots = (int32_t)slot_tiles.size();
  b->h_slot_seg = slot_seg;
  b->h_slot_tiles = slot_tiles;
  b->h_slot_phase = slot_phase;
  auto fail = [&](hipError_t e) {
    stgd_batch_free(b);
    return set_error(MMSB_ERR_HIP, "stgd batch: %s", hipGetErrorString(e));
  };
  hipError_t e;
  if ((e = hipMalloc(&b->d_ctx, sizeof(StepCtx) * ctxs.size())) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&b->d_segs, sizeof(BatchSeg) * segs.size())) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&b->d_slot_seg, sizeof(int32_t) * slot_seg.size())) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&b->d_slot_tiles, sizeof(int64_t) * slot_tiles.size())) != hipSuccess) return fail(e);
  if ((e = hipMemcpy(b->d_ctx, ctxs.data(), sizeof(StepCtx) * ctxs.size(), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e);
  if ((e = hipMemcpy(b->d_segs, segs.data(), sizeof(BatchSeg) * segs.size(), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e);
  if ((e = hipMemcpy(b->d_slot_seg, slot_seg.data(), sizeof(int32_t) * slot_seg.size(), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e);
  if ((e = hipMemcpy(b->d_slot_tiles, slot_tiles.data(), sizeof(int64_t) * slot_tiles.size(), hipMemcpyHostToDevice)) !=
      hipSuccess)
    return fail(e);
  int dev = 0, per_cu = 0, cus = 0;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return fail(e);
  if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel(precision), kThreads, 0)) != hipSuccess) return fail(e);
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return fail(e);
  if (per_cu < 1) {
    stgd_batch_free(b);
    return set_error(MMSB_ERR_HIP, "stgd batch: the persistent kernel does not fit on a compute unit");
  }
  b->grid_max = per_cu * cus;   // cooperative launch: every workgroup resident
  b->grid = cus;
  out = b;
  return MMSB_OK;
}

void stgd_batch_set_grid(StgdBatch* b, int32_t grid) {
  if (b && grid > 0) b->grid = std::min(grid, b->grid_max);
}

int32_t stgd_batch_grid(const StgdBatch* b) { return b ? b->grid : 0; }
int32_t stgd_batch_grid_max(const StgdBatch* b) { return b ? b->grid_max : 0; }

int stgd_batch_launch(hipStream_t s, StgdBatch* b) {
  MMSB_CHECK_ARG(b, "stgd batch: null plan");
  void* args[] = {&b->d_ctx, &b->d_segs, &b->d_slot_seg, &b->d_slot_tiles, &b->nslots};
  MMSB_HIP(hipLaunchCooperativeKernel(batch_kernel(b->precision), dim3(b->grid), dim3(kThreads), args, 0, s));
  return MMSB_OK;
}

template <int DMODE, int KMODE, typename TY, typename TPH>
static void launch_slot(hipStream_t s, const StgdBatch* b, int32_t sl) {
  const dim3 grid((unsigned)b->h_slot_tiles[sl]), block(kThreads);
  const int32_t s0 = b->h_slot_seg[sl], s1 = b->h_slot_seg[sl + 1];
  switch (b->h_slot_phase[sl]) {
#define MMSB_SLOT(PH)                                                                                              \
  case PH:                                                                                                         \
    hipLaunchKernelGGL((stgd_slot_phase_kernel<DMODE, KMODE, TY, TPH, PH>), grid, block, 0, s, b->d_ctx, b->d_segs, s0, s1); \
    break;
    MMSB_SLOT(PH_SETUP)
    MMSB_SLOT(PH_P1)
    MMSB_SLOT(PH_P2)
    MMSB_SLOT(PH_GRAM)
    MMSB_SLOT(PH_UPDATE)
    MMSB_SLOT(PH_COPY)
    MMSB_SLOT(PH_QPOLY)
    MMSB_SLOT(PH_FROB)
#undef MMSB_SLOT
    default:  // segments of different phases
      hipLaunchKernelGGL((stgd_slot_kernel<DMODE, KMODE, TY, TPH>), grid, block, 0, s, b->d_ctx, b->d_segs, s0, s1);
  }
}

int stgd_batch_launch_slots(hipStream_t s, StgdBatch* b) {
  MMSB_CHECK_ARG(b, "stgd batch: null plan");
  for (int32_t sl = 0; sl < b->nslots; ++sl) {
    if (b->precision == MMSB_PREC_BF16) launch_slot<MM_BF16, MM_BF16X3, __bf16, __bf16>(s, b, sl);
    else launch_slot<MM_F32, MM_F32, float, float>(s, b, sl);
    MMSB_LAUNCH_CHECK();
  }
  return MMSB_OK;
}

void stgd_batch_free(StgdBatch* b) {
  if (!b) return;
  (void)hipFree(b->d_ctx);
  (void)hipFree(b->d_segs);
  (void)hipFree(b->d_slot_seg);
  (void)hipFree(b->d_slot_tiles);
  delete b;
}

}  // namespace mmsb
