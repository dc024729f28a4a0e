// FP64 one-sided (Hestenes) Jacobi SVD of a small square matrix.
//
// Replaces the k x k culaDgesvd calls of the Nystrom helpers (gemm_pinv_nys.c:288,299) and
// serves both small factorisations of the whitening stage: the SVD of the r x r triangular factor
// of the randomized range of P_BC, and (after a positive shift) the eigendecomposition of the
// r x r Rayleigh-Ritz matrix of M2. Columns are orthogonalised pairwise by plane rotations
// A <- A J, V <- V J; at convergence sigma_j = ||A_j||, U_j = A_j / sigma_j.
//
// Blocks of B consecutive columns are paired by the round-robin (circle) schedule: nblk - 1 rounds
// of disjoint block pairs per sweep. One launch per round: the workgroup of block pair (P, Q) loads
// their 2B columns of A (and of V when it is accumulated) into LDS, orthogonalises every column pair
// among them (one inner sweep of 2B - 1 inner rounds of B disjoint pairs, one wave per pair), writes
// them back and adds its rotation count to the sweep's counter. After each sweep a one-thread kernel
// marks convergence (a sweep without rotations) and the host reads the mark. Plain launches keep the
// factorisation free of cooperative launches, whose global-wave-sync resource only one process on a
// GPU can hold, and a launch boundary costs less than a grid-wide barrier on this part.
#include "common.h"
#include "dense.h"

namespace mmsb {

static constexpr int kMaxSweeps = 40;
// Widest column block of the rotations (8, 4, 2 or 1); narrower blocks are taken when the wider do
// not fit in LDS.
#ifndef MMSB_JACOBI_BMAX
#define MMSB_JACOBI_BMAX 8
#endif
// counters: [0, kMaxSweeps) rotations per sweep, [kMaxSweeps] sweeps run, [kMaxSweeps + 1] converged
static constexpr int kSweepsRun = kMaxSweeps, kDone = kMaxSweeps + 1;

size_t jacobi_workspace_bytes(int32_t r) {
  return sizeof(double) * ((size_t)r * r + (size_t)r) + sizeof(int32_t) * (kMaxSweeps + 2) + 64;
}

__device__ inline double block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
  return s;
}

__device__ inline void pair_of(int rp, int round, int idx, int& p, int& q) {
  // circle method: player rp-1 fixed, the others rotate
  const int m = rp - 1;
  if (idx == 0) {
    p = m;
    q = round % m;
  } else {
    p = (round + idx) % m;
    q = (round - idx + m) % m;
  }
  if (p > q) {
    const int t = p;
    p = q;
    q = t;
  }
}

__global__ void jacobi_init_kernel(int32_t r, double* __restrict__ Vw, int32_t* __restrict__ counters) {
  const int64_t rr = (int64_t)r * r;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gsz = (int64_t)gridDim.x * blockDim.x;
  if (Vw)
    for (int64_t e = gtid; e < rr; e += gsz) Vw[e] = ((e / r) == (e % r)) ? 1.0 : 0.0;
  for (int64_t e = gtid; e < kMaxSweeps + 2; e += gsz) counters[e] = 0;
}

// One round of a sweep: workgroup b handles block pair b of the round (grid = block pairs).
template <int B, bool WV>
__global__ __launch_bounds__(64 * B) void jacobi_round_kernel(int32_t r, double* __restrict__ A, double* __restrict__ Vw,
                                                              int32_t* __restrict__ counters, int32_t sweep, int32_t round) {
  constexpr int TPG = 64;                  // one wave per column pair (64 B threads per workgroup)
  extern __shared__ double lds[];          // [2B][r] columns of A, then (WV) [2B][r] columns of V
  __shared__ int nrot;
  if (counters[kDone]) return;             // converged in an earlier sweep (uniform)
  double* la = lds;
  double* lv = lds + (int64_t)2 * B * r;
  const int nblk = (r + B - 1) / B;
  const int nbp = nblk + (nblk & 1);
  const int g = threadIdx.x / TPG, tg = threadIdx.x % TPG;
  int P, Q;
  pair_of(nbp, round, blockIdx.x, P, Q);
  auto gcol = [&](int lc) -> int {  // global column of local column lc, or -1
    const int blkid = lc < B ? P : Q;
    const int c = blkid * B + (lc % B);
    return (blkid < nblk && c < r) ? c : -1;
  };
  for (int lc = 0; lc < 2 * B; ++lc) {
    const int c = gcol(lc);
    if (c < 0) continue;
    for (int i = threadIdx.x; i < r; i += blockDim.x) {
      la[(int64_t)lc * r + i] = A[(int64_t)c * r + i];
      if constexpr (WV) lv[(int64_t)lc * r + i] = Vw[(int64_t)c * r + i];
    }
  }
  if (threadIdx.x == 0) nrot = 0;
  __syncthreads();
  int myrot = 0;
  for (int ir = 0; ir < 2 * B - 1; ++ir) {
    int lp, lq;
    pair_of(2 * B, ir, g, lp, lq);
    if (gcol(lp) >= 0 && gcol(lq) >= 0) {
      double* ap = la + (int64_t)lp * r;
      double* aq = la + (int64_t)lq * r;
      double a = 0.0, b = 0.0, gg = 0.0;
      for (int i = tg; i < r; i += TPG) {
        const double x = ap[i], y = aq[i];
        a = fma(x, x, a);
        b = fma(y, y, b);
        gg = fma(x, y, gg);
      }
      for (int off = TPG / 2; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
        gg += __shfl_xor(gg, off);
      }
      if (gg != 0.0 && fabs(gg) > 1e-15 * sqrt(a * b)) {
        const double zeta = (b - a) / (2.0 * gg);
        const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + t * t);
        const double sn = c * t;
        for (int i = tg; i < r; i += TPG) {
          const double x = ap[i], y = aq[i];
          ap[i] = c * x - sn * y;
          aq[i] = sn * x + c * y;
        }
        if constexpr (WV) {
          double* vp = lv + (int64_t)lp * r;
          double* vq = lv + (int64_t)lq * r;
          for (int i = tg; i < r; i += TPG) {
            const double x = vp[i], y = vq[i];
            vp[i] = c * x - sn * y;
            vq[i] = sn * x + c * y;
          }
        }
        if (tg == 0) ++myrot;
      }
    }
    __syncthreads();
  }
  if (tg == 0 && myrot) atomicAdd(&nrot, myrot);
  for (int lc = 0; lc < 2 * B; ++lc) {
    const int c = gcol(lc);
    if (c < 0) continue;
    for (int i = threadIdx.x; i < r; i += blockDim.x) {
      A[(int64_t)c * r + i] = la[(int64_t)lc * r + i];
      if constexpr (WV) Vw[(int64_t)c * r + i] = lv[(int64_t)lc * r + i];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && nrot) atomicAdd(counters + sweep, nrot);
}

__global__ void jacobi_sweep_end_kernel(int32_t* __restrict__ counters, int32_t sweep) {
  if (threadIdx.x != 0 || counters[kDone]) return;
  counters[kSweepsRun] = sweep + 1;
  if (counters[sweep] == 0) counters[kDone] = 1;
}

// sw[j] = ||A_j||, one workgroup per column.
__global__ __launch_bounds__(256) void jacobi_norms_kernel(int32_t r, const double* __restrict__ A, double* __restrict__ sw) {
  __shared__ double red[4];
  for (int j = blockIdx.x; j < r; j += gridDim.x) {
    const double* aj = A + (int64_t)j * r;
    double a = 0.0;
    for (int i = threadIdx.x; i < r; i += blockDim.x) a = fma(aj[i], aj[i], a);
    a = block_sum(a, red);
    if (threadIdx.x == 0) sw[j] = sqrt(a);
    __syncthreads();
  }
}

// Columns sorted by descending norm (stable): U = A_j / sigma_j, V, sigma; the sweep count.
__global__ __launch_bounds__(256) void jacobi_sort_kernel(int32_t r, const double* __restrict__ A, const double* __restrict__ Vw,
                                                          const double* __restrict__ sw, double* __restrict__ U,
                                                          double* __restrict__ sigma, double* __restrict__ V,
                                                          const int32_t* __restrict__ counters, int32_t* __restrict__ sweeps_out) {
  __shared__ double red[4];
  for (int j = blockIdx.x; j < r; j += gridDim.x) {
    const double sj = sw[j];
    int cnt = 0;
    for (int i = threadIdx.x; i < r; i += blockDim.x) {
      const double si = sw[i];
      cnt += (si > sj || (si == sj && i < j)) ? 1 : 0;
    }
    const int rank = (int)block_sum((double)cnt, red);
    const double inv = sj > 0.0 ? 1.0 / sj : 0.0;
    const double* aj = A + (int64_t)j * r;
    for (int i = threadIdx.x; i < r; i += blockDim.x) {
      U[(int64_t)rank * r + i] = aj[i] * inv;
      if (V) V[(int64_t)rank * r + i] = Vw[(int64_t)j * r + i];
    }
    if (threadIdx.x == 0) sigma[rank] = sj;
    __syncthreads();
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) sweeps_out[0] = counters[kSweepsRun];
}

template <int B, bool WV>
static int run_sweeps(hipStream_t s, int32_t r, double* A, double* Vw, int32_t* counters) {
  const void* fn = reinterpret_cast<const void*>(jacobi_round_kernel<B, WV>);
  const size_t shm = sizeof(double) * (size_t)(WV ? 4 : 2) * B * r;
  MMSB_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
  const int nblk = (r + B - 1) / B;
  const int nbp = nblk + (nblk & 1);
  const int npairs = nbp / 2;
  int32_t done = 0;
  for (int sweep = 0; sweep < kMaxSweeps && !done; ++sweep) {
    for (int round = 0; round < nbp - 1; ++round) {
      hipLaunchKernelGGL((jacobi_round_kernel<B, WV>), dim3(npairs), dim3(64 * B), shm, s, r, A, Vw, counters, sweep,
                         round);
      MMSB_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(jacobi_sweep_end_kernel, dim3(1), dim3(64), 0, s, counters, sweep);
    MMSB_LAUNCH_CHECK();
    MMSB_HIP(hipMemcpyAsync(&done, counters + kDone, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    MMSB_HIP(hipStreamSynchronize(s));
  }
  if (!done)  // every sweep still rotated: the factors are not converged, so say so instead of using them
    return set_error(MMSB_ERR_NUMERIC, "jacobi: no convergence after %d sweeps (r = %d)", kMaxSweeps, r);
  return MMSB_OK;
}

int jacobi_svd(hipStream_t s, int32_t r, double* A, double* U, double* sigma, double* V, void* work, int32_t* sweeps_out) {
  MMSB_CHECK_ARG(r >= 1 && r <= 8192, "jacobi: r=%d out of range", r);
  double* Vw = static_cast<double*>(work);
  double* sw = Vw + (size_t)r * r;
  int32_t* counters = reinterpret_cast<int32_t*>(sw + r);
  // 8-column blocks while 2B columns of A (and of V when it is wanted) fit in the 160 KiB of LDS;
  // without V the rotations touch half the data. (16-column blocks, which V-free would allow at
  // r = 510, were slower: the longer inner rounds outweighed the halved round count.)
  auto fits = [&](int b, int arrays) { return sizeof(double) * (size_t)2 * arrays * b * r + 256 <= (size_t)160 * 1024; };
  const int grid = (int)std::min<int64_t>(std::max<int64_t>(1, ((int64_t)r * r + 255) / 256), 1024);
  hipLaunchKernelGGL(jacobi_init_kernel, dim3(grid), dim3(256), 0, s, r, V ? Vw : nullptr, counters);
  MMSB_LAUNCH_CHECK();
  const int arrays = V ? 2 : 1;
  int st;
  if (MMSB_JACOBI_BMAX >= 8 && fits(8, arrays))
    st = V ? run_sweeps<8, true>(s, r, A, Vw, counters) : run_sweeps<8, false>(s, r, A, Vw, counters);
  else if (MMSB_JACOBI_BMAX >= 4 && fits(4, arrays))
    st = V ? run_sweeps<4, true>(s, r, A, Vw, counters) : run_sweeps<4, false>(s, r, A, Vw, counters);
  else if (MMSB_JACOBI_BMAX >= 2 && fits(2, arrays))
    st = V ? run_sweeps<2, true>(s, r, A, Vw, counters) : run_sweeps<2, false>(s, r, A, Vw, counters);
  else if (fits(1, arrays))
    st = V ? run_sweeps<1, true>(s, r, A, Vw, counters) : run_sweeps<1, false>(s, r, A, Vw, counters);
  else
    return set_error(MMSB_ERR_ARG, "jacobi: r=%d does not fit the LDS of one workgroup", r);
  MMSB_TRY(st);
  const int cols = std::min(r, 1024);
  hipLaunchKernelGGL(jacobi_norms_kernel, dim3(cols), dim3(256), 0, s, r, A, sw);
  MMSB_LAUNCH_CHECK();
  hipLaunchKernelGGL(jacobi_sort_kernel, dim3(cols), dim3(256), 0, s, r, A, Vw, sw, U, sigma, V, counters, sweeps_out);
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

}  // namespace mmsb
