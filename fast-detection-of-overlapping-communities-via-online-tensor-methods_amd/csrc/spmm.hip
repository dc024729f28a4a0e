// CSR x dense SpMM for gfx950: Y = alpha * G * B, FP32, row-major B and Y.
//
// Every sparse-times-dense product of the pipeline goes through this kernel: the randomized
// range finder of P_BC = G_XB^T G_XC / n_X, the Z_B / Z_C factors, the whitened neighbour
// vectors y_A, y_B, y_C and the four Pi products. The reference forms these as dense n x n
// DGEMMs (culaBlockDgemm, gemm_pinv_nys.c:27-145; pre_proc.c:154-366; post_proc.c:87-106);
// here the adjacency blocks stay in CSR and B is a tall n x r matrix.
//
// Mapping: one wave64 per output row and 256-column strip of it, rows taken in the order of an
// optional permutation (the pipeline passes rows sorted by descending length, so that the longest
// rows -- hub nodes with thousands of nonzeros on degree-corrected graphs -- start first instead of
// forming the tail). The row's column indices are read 64 at a time with one
// coalesced load and broadcast with ds_bpermute (__shfl). Each lane owns float4 column chunks of
// the dense rows it gathers, so one gather of a B row is a single contiguous 16*64-byte access
// per wave instruction. When r/4 < 64 the wave is split into G = 64/LPG lane groups that gather
// G different nonzeros at once, then reduced across groups with butterfly shuffles. The nonzero
// loop is unrolled so that several independent B-row gathers are in flight per wave -- the
// kernel is HBM-bound (each nonzero moves 4r bytes of B) and needs that memory-level
// parallelism to approach the bandwidth roofline.
#include <atomic>

#include "common.h"
#include "dense.h"


namespace mmsb {

// Largest gathered matrix B (bytes) for which the whole-matrix and minibatch-rows SpMMs cut wide
// rows into 256-column strips: 8 GiB unless mmsb_set_strip_limit sets another (process-wide).
static constexpr int64_t kStripBytesDefault = (int64_t)8 << 30;
static std::atomic<int64_t> g_strip_bytes{kStripBytesDefault};

// Gathers one output row: acc[p] (on lanes of group 0) = sum over the row's nonzeros of v * B[c, :],
// float4 chunk q = gl + p * LPG. Shared by the whole-matrix kernel and the minibatch-rows kernel, so
// that a row is summed in the same order by both.
template <int LPG, int P, bool VALS>
__device__ __forceinline__ void row_gather(floatx4 (&acc)[P], int64_t row, const int32_t* __restrict__ rowptr,
                                           const int32_t* __restrict__ colidx, const float* __restrict__ vals,
                                           const float* __restrict__ B, int32_t ldb, int32_t r4, int lane) {
  constexpr int G = 64 / LPG;  // nonzeros gathered concurrently by one wave
  const int grp = lane / LPG;
  const int gl = lane % LPG;
#pragma unroll
  for (int p = 0; p < P; ++p) acc[p] = floatx4{0.f, 0.f, 0.f, 0.f};

  const int32_t beg = rowptr[row];
  const int32_t end = rowptr[row + 1];
  for (int32_t base = beg; base < end; base += 64) {
    const int32_t cnt = min(64, end - base);
    const int32_t my_c = (lane < cnt) ? colidx[base + lane] : 0;
    float my_v = 1.f;
    if (VALS) my_v = (lane < cnt) ? vals[base + lane] : 0.f;
    // Unrolled by U groups of G nonzeros: U*P independent 16-byte gathers in flight per lane.
    constexpr int U = (P >= 8) ? 1 : 8 / P;
    int32_t j = 0;
    for (; j + U * G <= cnt; j += U * G) {
      int32_t c[U];
      float v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        c[u] = __shfl(my_c, j + u * G + grp);
        v[u] = VALS ? __shfl(my_v, j + u * G + grp) : 1.f;
      }
      floatx4 x[U][P];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const floatx4* brow = reinterpret_cast<const floatx4*>(B + (int64_t)c[u] * ldb);
#pragma unroll
        for (int p = 0; p < P; ++p) {
          const int q = gl + p * LPG;
          x[u][p] = (q < r4) ? brow[q] : floatx4{0.f, 0.f, 0.f, 0.f};
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int p = 0; p < P; ++p) acc[p] += VALS ? v[u] * x[u][p] : x[u][p];
    }
    for (; j < cnt; j += G) {
      const int32_t jj = j + grp;
      const int32_t c = __shfl(my_c, jj < 64 ? jj : 63);
      const float v = VALS ? __shfl(my_v, jj < 64 ? jj : 63) : 1.f;
      if (jj < cnt) {
        const floatx4* brow = reinterpret_cast<const floatx4*>(B + (int64_t)c * ldb);
#pragma unroll
        for (int p = 0; p < P; ++p) {
          const int q = gl + p * LPG;
          if (q < r4) acc[p] += VALS ? v * brow[q] : brow[q];
        }
      }
    }
  }
  // Reduce the G lane groups (lanes gl, gl+LPG, gl+2*LPG, ...) onto group 0.
#pragma unroll
  for (int off = LPG; off < 64; off <<= 1) {
#pragma unroll
    for (int p = 0; p < P; ++p) {
      acc[p].x += __shfl_xor(acc[p].x, off);
      acc[p].y += __shfl_xor(acc[p].y, off);
      acc[p].z += __shfl_xor(acc[p].z, off);
      acc[p].w += __shfl_xor(acc[p].w, off);
    }
  }
}

template <int LPG, int P, bool VALS>
__global__ __launch_bounds__(256) void spmm_csr_rowwave(int32_t m, const int32_t* __restrict__ rowptr,
                                                        const int32_t* __restrict__ colidx,
                                                        const float* __restrict__ vals, float alpha,
                                                        const float* __restrict__ B, int32_t ldb,
                                                        int32_t r, float* __restrict__ Y, int32_t ldy,
                                                        const int32_t* __restrict__ order) {
  // column strip blockIdx.y of 4 * LPG * P columns (one strip unless r is wider)
  const int32_t col0 = (int32_t)blockIdx.y * (4 * LPG * P);
  B += col0;
  Y += col0;
  r = min(r - col0, 4 * LPG * P);
  const int32_t r4 = (r + 3) >> 2;
  const int lane = threadIdx.x & 63;
  const int64_t slot = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (slot >= m) return;  // uniform per wave
  const int64_t row = order ? order[slot] : slot;
  const int grp = lane / LPG;
  const int gl = lane % LPG;
  floatx4 acc[P];
  row_gather<LPG, P, VALS>(acc, row, rowptr, colidx, vals, B, ldb, r4, lane);
  if (grp == 0) {
    float* yrow = Y + row * (int64_t)ldy;
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int q = gl + p * LPG;
      const floatx4 o = alpha * acc[p];
      if (4 * q + 4 <= r) {
        __builtin_nontemporal_store(o, reinterpret_cast<floatx4*>(yrow) + q);
      } else if (q < r4) {  // last partial chunk: columns >= r are left untouched
        yrow[4 * q] = o.x;
        if (4 * q + 1 < r) yrow[4 * q + 1] = o.y;
        if (4 * q + 2 < r) yrow[4 * q + 2] = o.z;
      }
    }
  }
}

template <int LPG, int P>
static int launch_spmm(hipStream_t s, int32_t m, const int32_t* rowptr, const int32_t* colidx,
                       const float* vals, float alpha, const float* B, int32_t ldb, int32_t r,
                       float* Y, int32_t ldy, const int32_t* order) {
  const int waves_per_block = 4;
  const dim3 grid((unsigned)ceil_div(m, waves_per_block), (unsigned)ceil_div(r, 4 * LPG * P));
  if (vals)
    hipLaunchKernelGGL((spmm_csr_rowwave<LPG, P, true>), grid, dim3(64 * waves_per_block), 0, s, m,
                       rowptr, colidx, vals, alpha, B, ldb, r, Y, ldy, order);
  else
    hipLaunchKernelGGL((spmm_csr_rowwave<LPG, P, false>), grid, dim3(64 * waves_per_block), 0, s, m,
                       rowptr, colidx, vals, alpha, B, ldb, r, Y, ldy, order);
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

int spmm_f32(hipStream_t s, int32_t m, int32_t n, int64_t nnz, const int32_t* rowptr,
             const int32_t* colidx, const float* vals, float alpha, const float* B, int32_t ldb,
             int32_t r, float* Y, int32_t ldy, const int32_t* order) {
  MMSB_CHECK_ARG(m >= 0 && n >= 0 && nnz >= 0 && nnz < (int64_t)INT32_MAX, "spmm: bad sizes m=%d n=%d nnz=%lld", m, n, (long long)nnz);
  MMSB_CHECK_ARG(r >= 1 && r <= ldb && r <= ldy, "spmm: need 1 <= r <= ldb, ldy (r=%d ldb=%d ldy=%d)", r, ldb, ldy);
  MMSB_CHECK_ARG(ldb % 4 == 0 && ldy % 4 == 0, "spmm: ldb and ldy must be multiples of 4");
  MMSB_CHECK_ARG(r <= 4096, "spmm: r=%d above the supported 4096", r);
  if (m == 0) return MMSB_OK;
  MMSB_CHECK_ARG(rowptr && Y && (nnz == 0 || (colidx && B)), "spmm: null pointer");
  MMSB_CHECK_ARG(aligned16(B) && aligned16(Y), "spmm: B and Y must be 16-byte aligned");
  const int32_t r4 = (r + 3) / 4;
  if (r4 <= 1) return launch_spmm<1, 1>(s, m, rowptr, colidx, vals, alpha, B, ldb, r, Y, ldy, order);
  if (r4 <= 2) return launch_spmm<2, 1>(s, m, rowptr, colidx, vals, alpha, B, ldb, r, Y, ldy, order);
  if (r4 <= 4) return launch_spmm<4, 1>(s, m, rowptr, colidx, vals, alpha, B, ldb, r, Y, ldy, order);
  if (r4 <= 8) return launch_spmm<8, 1>(s, m, rowptr, colidx, vals, alpha, B, ldb, r, Y, ldy, order);
  if (r4 <= 16) return launch_spmm<16, 1>(s, m, rowptr, colidx, vals, alpha, B, ldb, r, Y, ldy, order);
  if (r4 <= 32) return launch_spmm<32, 1>(s, m, rowptr, colidx, vals, alpha, B, ldb, r, Y, ldy, order);
  // Wider than 256 columns and B within the strip limit: 256-column strips, one wave per (row, strip)
  // (grid.y = strips), twice or four times the waves of one wave per row with 2-4 float4 chunks per
  // lane (1M x 1M blocks: r = 510 7.7 -> 6.9 ms, r = 1010 16.5 -> 15.7 ms). On a larger B each strip
  // pass touches every row's pages again, and the strips lose (4M rows, r = 1010: 129.5 -> 137.4
  // ms), so there a wave keeps the whole row. Every form sums a row's nonzeros in the same order
  // (LPG = 64: one nonzero after another), so Y does not depend on the choice.
  if ((int64_t)n * ldb * (int64_t)sizeof(float) <= g_strip_bytes.load(std::memory_order_relaxed))
    return launch_spmm<64, 1>(s, m, rowptr, colidx, vals, alpha, B, ldb, r, Y, ldy, order);
  if (r4 <= 128) return launch_spmm<64, 2>(s, m, rowptr, colidx, vals, alpha, B, ldb, r, Y, ldy, order);
  if (r4 <= 192) return launch_spmm<64, 3>(s, m, rowptr, colidx, vals, alpha, B, ldb, r, Y, ldy, order);
  if (r4 <= 256) return launch_spmm<64, 4>(s, m, rowptr, colidx, vals, alpha, B, ldb, r, Y, ldy, order);
  if (r4 <= 512) return launch_spmm<64, 8>(s, m, rowptr, colidx, vals, alpha, B, ldb, r, Y, ldy, order);
  return launch_spmm<64, 16>(s, m, rowptr, colidx, vals, alpha, B, ldb, r, Y, ldy, order);
}

// ---- whitened rows of one STGD minibatch -----------------------------------------------------------
// Recomputes y_A, y_B, y_C of rows [row0, row0 + nb) of X from the CSR blocks (y = G_X* rows times
// W, Z_B^T W, Z_C^T W; pre_proc.c:339-345 forms all of them up front) for STGD runs that do not keep
// the n_X x k neighbour vectors resident. blockIdx.y selects the block; each row is summed by
// row_gather exactly as the whole-matrix kernel sums it, and converted to the output type by the same
// rounding as the stored vectors, so both forms feed STGD identical operands.
template <int LPG, int P, typename TO>
__global__ __launch_bounds__(256) void yrows_kernel(YRowsJobs jobs, int32_t nb, int32_t r, int32_t ldr, float alpha) {
  // blockIdx.y = job * strips + strip: column strips of 4 * LPG * P columns, as in spmm_csr_rowwave
  const int strips = (int)(gridDim.y / (unsigned)jobs.count);
  const int job = (int)blockIdx.y / strips, strip = (int)blockIdx.y % strips;
  YRowsJob J = job == 0 ? jobs.j[0] : (job == 1 ? jobs.j[1] : jobs.j[2]);
  const int32_t col0 = strip * (4 * LPG * P);
  J.R += col0;
  r = min(r - col0, 4 * LPG * P);
  const int32_t r4 = (r + 3) >> 2;
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (row >= nb) return;  // uniform per wave
  const int grp = lane / LPG;
  const int gl = lane % LPG;
  floatx4 acc[P];
  if (J.vals)
    row_gather<LPG, P, true>(acc, row, J.rowptr, J.colidx, J.vals, J.R, ldr, r4, lane);
  else
    row_gather<LPG, P, false>(acc, row, J.rowptr, J.colidx, nullptr, J.R, ldr, r4, lane);
  if (grp != 0) return;
  TO* out = static_cast<TO*>(J.out) + (J.trans ? (int64_t)col0 * J.ldo : (int64_t)col0);
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int q = gl + p * LPG;
    if (q >= r4) continue;
    const floatx4 o = alpha * acc[p];
    const TO v[4] = {(TO)o.x, (TO)o.y, (TO)o.z, (TO)o.w};
    if (J.trans) {  // [r x ldo]: column `row` of the transposed block
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (4 * q + e < r) out[(int64_t)(4 * q + e) * J.ldo + row] = v[e];
    } else {
      TO* yrow = out + row * J.ldo;
      if (4 * q + 4 <= r) {  // whole chunk: one 16-byte (FP32) or 8-byte (BF16) store
        if constexpr (sizeof(TO) == 4) {
          *reinterpret_cast<floatx4*>(yrow + 4 * q) = o;
        } else {
          typedef TO tox4 __attribute__((ext_vector_type(4)));
          *reinterpret_cast<tox4*>(yrow + 4 * q) = tox4{v[0], v[1], v[2], v[3]};
        }
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (4 * q + e < r) yrow[4 * q + e] = v[e];
      }
    }
  }
}

template <int LPG, int P>
static int launch_yrows(hipStream_t s, const YRowsJobs& jobs, int32_t nb, int32_t r, int32_t ldr, bool bf16) {
  const int strips = (int)ceil_div(r, 4 * LPG * P);
  const dim3 grid((unsigned)ceil_div(nb, 4), (unsigned)(jobs.count * strips)), block(256);
  if (bf16)
    hipLaunchKernelGGL((yrows_kernel<LPG, P, __bf16>), grid, block, 0, s, jobs, nb, r, ldr, 1.f);
  else
    hipLaunchKernelGGL((yrows_kernel<LPG, P, float>), grid, block, 0, s, jobs, nb, r, ldr, 1.f);
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

int spmm_yrows(hipStream_t s, const YRowsJobs& jobs, int32_t nb, int32_t r, int32_t ldr, bool bf16) {
  MMSB_CHECK_ARG(nb >= 0 && r >= 1 && r <= ldr && r <= 4096 && ldr % 4 == 0, "yrows: bad sizes nb=%d r=%d ldr=%d", nb, r, ldr);
  MMSB_CHECK_ARG(jobs.count >= 1 && jobs.count <= 3, "yrows: 1 to 3 jobs (got %d)", jobs.count);
  if (nb == 0) return MMSB_OK;
  for (int32_t i = 0; i < jobs.count; ++i) {
    const YRowsJob& j = jobs.j[i];
    MMSB_CHECK_ARG(j.rowptr && j.colidx && j.R && j.out, "yrows: null pointer");
    MMSB_CHECK_ARG(aligned16(j.R), "yrows: R must be 16-byte aligned");
    MMSB_CHECK_ARG(j.trans ? j.ldo >= nb : j.ldo >= r, "yrows: output leading dimension %lld too small", (long long)j.ldo);
    MMSB_CHECK_ARG(j.trans || (j.ldo % 4 == 0 && aligned16(j.out)), "yrows: row-major output needs ldo %% 4 == 0 and 16-byte alignment");
  }
  const int32_t r4 = (r + 3) / 4;
  if (r4 <= 1) return launch_yrows<1, 1>(s, jobs, nb, r, ldr, bf16);
  if (r4 <= 2) return launch_yrows<2, 1>(s, jobs, nb, r, ldr, bf16);
  if (r4 <= 4) return launch_yrows<4, 1>(s, jobs, nb, r, ldr, bf16);
  if (r4 <= 8) return launch_yrows<8, 1>(s, jobs, nb, r, ldr, bf16);
  if (r4 <= 16) return launch_yrows<16, 1>(s, jobs, nb, r, ldr, bf16);
  if (r4 <= 32) return launch_yrows<32, 1>(s, jobs, nb, r, ldr, bf16);
  int64_t rmax = 0;  // the largest gathered matrix decides, as in spmm_f32
  for (int32_t i = 0; i < jobs.count; ++i) rmax = std::max<int64_t>(rmax, jobs.j[i].rows_R);
  if (rmax * ldr * (int64_t)sizeof(float) <= g_strip_bytes.load(std::memory_order_relaxed)) return launch_yrows<64, 1>(s, jobs, nb, r, ldr, bf16);
  if (r4 <= 128) return launch_yrows<64, 2>(s, jobs, nb, r, ldr, bf16);
  if (r4 <= 192) return launch_yrows<64, 3>(s, jobs, nb, r, ldr, bf16);
  if (r4 <= 256) return launch_yrows<64, 4>(s, jobs, nb, r, ldr, bf16);
  if (r4 <= 512) return launch_yrows<64, 8>(s, jobs, nb, r, ldr, bf16);
  return launch_yrows<64, 16>(s, jobs, nb, r, ldr, bf16);
}

}  // namespace mmsb

extern "C" int mmsb_spmm_f32(void* stream, int32_t m, int32_t n, int64_t nnz, const int32_t* rowptr,
                             const int32_t* colidx, const float* vals, float alpha, const float* B,
                             int32_t ldb, int32_t r, float* Y, int32_t ldy) {
  mmsb::clear_error();
  return mmsb::spmm_f32(static_cast<hipStream_t>(stream), m, n, nnz, rowptr, colidx, vals, alpha, B,
                        ldb, r, Y, ldy, nullptr);
}

extern "C" int mmsb_spmm_f32_ordered(void* stream, int32_t m, int32_t n, int64_t nnz, const int32_t* rowptr,
                                     const int32_t* colidx, const float* vals, float alpha, const float* B,
                                     int32_t ldb, int32_t r, float* Y, int32_t ldy, const int32_t* row_order) {
  mmsb::clear_error();
  MMSB_CHECK_ARG(row_order, "spmm_ordered: null row order");
  return mmsb::spmm_f32(static_cast<hipStream_t>(stream), m, n, nnz, rowptr, colidx, vals, alpha, B,
                        ldb, r, Y, ldy, row_order);
}

extern "C" int64_t mmsb_set_strip_limit(int64_t bytes) {
  const int64_t v = bytes < 0 ? mmsb::kStripBytesDefault : bytes;
  return mmsb::g_strip_bytes.exchange(v);
}
