// FP64 one-sided (Hestenes) Jacobi SVD of a small square matrix, one cooperative launch.
//
// Replaces the k x k culaDgesvd calls of the Nystrom helpers (gemm_pinv_nys.c:288,299) and
// serves both small factorisations of the whitening stage: the SVD of the r x r triangular factor
// of the randomized range of P_BC, and (after a positive shift) the eigendecomposition of the
// r x r Rayleigh-Ritz matrix of M2. Columns are orthogonalised pairwise by plane rotations
// A <- A J, V <- V J; at convergence sigma_j = ||A_j||, U_j = A_j / sigma_j.
//
// Parallel ordering: the round-robin (circle) schedule pairs the r columns into r/2 disjoint
// pairs per round, r-1 rounds per sweep. Workgroups own pairs; a grid-wide barrier
// (cooperative launch, so every workgroup is resident) separates rounds. Each sweep counts its
// rotations in its own counter; all workgroups read the count after the sweep's last barrier and
// leave the loop together when it is zero or after kMaxSweeps sweeps, so every wave exits.
#include <hip/hip_cooperative_groups.h>

#include "common.h"
#include "dense.h"

namespace cg = cooperative_groups;

namespace mmsb {

static constexpr int kMaxSweeps = 40;
static constexpr int kJacobiThreads = 256;

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

__global__ __launch_bounds__(kJacobiThreads) void jacobi_svd_kernel(int32_t r, double* __restrict__ A,
                                                                     double* __restrict__ U, double* __restrict__ sigma,
                                                                     double* __restrict__ V, double* __restrict__ Vw,
                                                                     double* __restrict__ sw, int32_t* __restrict__ counters,
                                                                     int32_t* __restrict__ sweeps_out) {
  cg::grid_group grid = cg::this_grid();
  __shared__ double red[kJacobiThreads / 64];
  __shared__ int go_on;
  const int nwg = gridDim.x;
  const int64_t rr = (int64_t)r * r;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gsz = (int64_t)nwg * blockDim.x;
  for (int64_t e = gtid; e < rr; e += gsz) Vw[e] = ((e / r) == (e % r)) ? 1.0 : 0.0;
  for (int64_t e = gtid; e < kMaxSweeps + 2; e += gsz) counters[e] = 0;
  grid.sync();

  const int rp = r + (r & 1);
  const int npairs = rp / 2;
  int sweep = 0;
  for (; sweep < kMaxSweeps; ++sweep) {
    for (int round = 0; round < rp - 1; ++round) {
      for (int pi = blockIdx.x; pi < npairs; pi += nwg) {
        int p, q;
        pair_of(rp, round, pi, p, q);
        if (q >= r) continue;  // pair with the padding column: no-op (uniform per workgroup)
        double* ap = A + (int64_t)p * r;
        double* aq = A + (int64_t)q * r;
        double a = 0.0, b = 0.0, g = 0.0;
        for (int i = threadIdx.x; i < r; i += blockDim.x) {
          const double x = ap[i], y = aq[i];
          a = fma(x, x, a);
          b = fma(y, y, b);
          g = fma(x, y, g);
        }
        a = block_sum(a, red);
        b = block_sum(b, red);
        g = block_sum(g, red);
        if (g != 0.0 && fabs(g) > 1e-15 * sqrt(a * b)) {
          const double zeta = (b - a) / (2.0 * g);
          const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double c = 1.0 / sqrt(1.0 + t * t);
          const double s = c * t;
          for (int i = threadIdx.x; i < r; i += blockDim.x) {
            const double x = ap[i], y = aq[i];
            ap[i] = c * x - s * y;
            aq[i] = s * x + c * y;
          }
          double* vp = Vw + (int64_t)p * r;
          double* vq = Vw + (int64_t)q * r;
          for (int i = threadIdx.x; i < r; i += blockDim.x) {
            const double x = vp[i], y = vq[i];
            vp[i] = c * x - s * y;
            vq[i] = s * x + c * y;
          }
          if (threadIdx.x == 0) atomicAdd(counters + sweep, 1);
        }
        __syncthreads();
      }
      grid.sync();
    }
    if (threadIdx.x == 0) go_on = atomicAdd(counters + sweep, 0) != 0;
    __syncthreads();
    if (!go_on) break;
  }
  // column norms
  for (int j = blockIdx.x; j < r; j += nwg) {
    const double* aj = A + (int64_t)j * r;
    double a = 0.0;
    for (int i = threadIdx.x; i < r; i += blockDim.x) a = fma(aj[i], aj[i], a);
    a = block_sum(a, red);
    if (threadIdx.x == 0) sw[j] = sqrt(a);
    __syncthreads();
  }
  grid.sync();
  // sort descending (stable) and write U, sigma, V
  for (int j = blockIdx.x; j < r; j += nwg) {
    const double sj = sw[j];
    int cnt = 0;
    for (int i = threadIdx.x; i < r; i += blockDim.x) {
      const double si = sw[i];
      cnt += (si > sj || (si == sj && i < j)) ? 1 : 0;
    }
    const int rank = (int)block_sum((double)cnt, red);
    const double inv = sj > 0.0 ? 1.0 / sj : 0.0;
    const double* aj = A + (int64_t)j * r;
    const double* vj = Vw + (int64_t)j * r;
    for (int i = threadIdx.x; i < r; i += blockDim.x) {
      U[(int64_t)rank * r + i] = aj[i] * inv;
      if (V) V[(int64_t)rank * r + i] = vj[i];
    }
    if (threadIdx.x == 0) sigma[rank] = sj;
    __syncthreads();
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) sweeps_out[0] = sweep + (sweep < kMaxSweeps ? 1 : 0);
}

// Block variant: blocks of B consecutive columns; a round pairs blocks by the same circle
// schedule, and the workgroup holding blocks (P, Q) loads their 2B columns of A and of V into LDS,
// orthogonalises every pair among them (one inner sweep, 2B - 1 inner rounds of B disjoint pairs,
// 256 / B threads per pair), and writes them back. A sweep needs nblk - 1 grid barriers instead
// of r - 1. Convergence test, ordering of the output and exit protocol as above.
template <int B, bool WV>
__global__ __launch_bounds__(64 * B) void jacobi_block_kernel(int32_t r, double* __restrict__ A,
                                                                       double* __restrict__ U, double* __restrict__ sigma,
                                                                       double* __restrict__ V, double* __restrict__ Vw,
                                                                       double* __restrict__ sw, int32_t* __restrict__ counters,
                                                                       int32_t* __restrict__ sweeps_out) {
  constexpr int TPG = 64;                  // one wave per column pair (64 B threads per workgroup)
  extern __shared__ double lds[];          // [2B][r] columns of A, then (WV) [2B][r] columns of V
  cg::grid_group grid = cg::this_grid();
  __shared__ double red[B];
  __shared__ int go_on, nrot;
  double* la = lds;
  double* lv = lds + (int64_t)2 * B * r;
  const int nwg = gridDim.x;
  const int64_t rr = (int64_t)r * r;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gsz = (int64_t)nwg * blockDim.x;
  if constexpr (WV)
    for (int64_t e = gtid; e < rr; e += gsz) Vw[e] = ((e / r) == (e % r)) ? 1.0 : 0.0;
  for (int64_t e = gtid; e < kMaxSweeps + 2; e += gsz) counters[e] = 0;
  grid.sync();

  const int nblk = (r + B - 1) / B;
  const int nbp = nblk + (nblk & 1);
  const int npairs = nbp / 2;
  const int g = threadIdx.x / TPG, tg = threadIdx.x % TPG;
  int sweep = 0;
  for (; sweep < kMaxSweeps; ++sweep) {
    for (int round = 0; round < nbp - 1; ++round) {
      for (int pi = blockIdx.x; pi < npairs; pi += nwg) {
        int P, Q;
        pair_of(nbp, round, pi, P, Q);
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
        __syncthreads();
      }
      grid.sync();
    }
    if (threadIdx.x == 0) go_on = atomicAdd(counters + sweep, 0) != 0;
    __syncthreads();
    if (!go_on) break;
  }
  for (int j = blockIdx.x; j < r; j += nwg) {
    const double* aj = A + (int64_t)j * r;
    double a = 0.0;
    for (int i = threadIdx.x; i < r; i += blockDim.x) a = fma(aj[i], aj[i], a);
    a = block_sum(a, red);
    if (threadIdx.x == 0) sw[j] = sqrt(a);
    __syncthreads();
  }
  grid.sync();
  for (int j = blockIdx.x; j < r; j += nwg) {
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
      if constexpr (WV) V[(int64_t)rank * r + i] = Vw[(int64_t)j * r + i];
    }
    if (threadIdx.x == 0) sigma[rank] = sj;
    __syncthreads();
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) sweeps_out[0] = sweep + (sweep < kMaxSweeps ? 1 : 0);
}

template <int B, bool WV>
static int launch_block(hipStream_t s, int32_t r, double* A, double* U, double* sigma, double* V, double* Vw, double* sw,
                        int32_t* counters, int32_t* sweeps_out) {
  const void* fn = reinterpret_cast<const void*>(jacobi_block_kernel<B, WV>);
  const size_t shm = sizeof(double) * (size_t)(WV ? 4 : 2) * B * r;
  MMSB_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
  int dev = 0, per_cu = 0, cus = 0;
  MMSB_HIP(hipGetDevice(&dev));
  MMSB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64 * B, shm));
  MMSB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  if (per_cu < 1) return set_error(MMSB_ERR_HIP, "jacobi: block kernel does not fit (%zu bytes of LDS)", shm);
  const int nblk = (r + B - 1) / B;
  const int npairs = (nblk + (nblk & 1)) / 2;
  const int nwg = std::max(1, std::min(npairs, per_cu * cus));
  void* args[] = {&r, &A, &U, &sigma, &V, &Vw, &sw, &counters, &sweeps_out};
  MMSB_HIP(hipLaunchCooperativeKernel(fn, dim3(nwg), dim3(64 * B), args, shm, s));
  return MMSB_OK;
}

int jacobi_svd(hipStream_t s, int32_t r, double* A, double* U, double* sigma, double* V, void* work, int32_t* sweeps_out) {
  MMSB_CHECK_ARG(r >= 1 && r <= 8192, "jacobi: r=%d out of range", r);
  double* Vw = static_cast<double*>(work);
  double* sw = Vw + (size_t)r * r;
  int32_t* counters = reinterpret_cast<int32_t*>(sw + r);
  // block variant while 2B columns of A (and of V when it is wanted) fit in the 160 KiB of LDS. Without
  // V the rotations touch half the data. (16-column blocks, which V-free would allow at r = 510, were
  // slower: 51.8 against 48 ms, the longer inner rounds outweighing the halved grid barriers.)
  auto fits = [&](int b, int arrays) { return sizeof(double) * (size_t)2 * arrays * b * r + 256 <= (size_t)160 * 1024; };
  if (!V && r >= 64 && fits(8, 1)) return launch_block<8, false>(s, r, A, U, sigma, V, Vw, sw, counters, sweeps_out);
  if (r >= 64 && fits(8, 2)) return launch_block<8, true>(s, r, A, U, sigma, V, Vw, sw, counters, sweeps_out);
  if (r >= 64 && fits(4, 2)) return launch_block<4, true>(s, r, A, U, sigma, V, Vw, sw, counters, sweeps_out);
  int dev = 0;
  MMSB_HIP(hipGetDevice(&dev));
  int per_cu = 0, cus = 0;
  MMSB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(jacobi_svd_kernel),
                                                        kJacobiThreads, 0));
  MMSB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int max_resident = std::max(1, per_cu * cus);
  const int npairs = (r + (r & 1)) / 2;
  const int nwg = std::max(1, std::min(npairs, std::min(max_resident, cus)));
  void* args[] = {&r, &A, &U, &sigma, &V, &Vw, &sw, &counters, &sweeps_out};
  MMSB_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(jacobi_svd_kernel), dim3(nwg), dim3(kJacobiThreads),
                                      args, 0, s));
  return MMSB_OK;
}

}  // namespace mmsb
