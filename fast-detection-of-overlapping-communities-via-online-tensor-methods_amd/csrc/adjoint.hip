// Loss and exact adjoint of the implicit tensor objective with respect to Phi.
//
//   L(Phi) = theta/2 sum_ij <phi_i, phi_j>^3 - sum_i T_b(phi_i, phi_i, phi_i)
//   T_b(u,v,w) = mean_x [c3 (y_A.u)(y_B.v)(y_C.w) + c1 ((y_A.u)(y_B.v)(mu_C.w) + (y_A.u)(mu_B.v)(y_C.w)
//                + (mu_A.u)(y_B.v)(y_C.w))] + c0 (mu_A.u)(mu_B.v)(mu_C.w)
//   dL/dPhi = 3 theta Phi (G o G) - [T(I,phi,phi) + T(phi,I,phi) + T(phi,phi,I)]   (G = Phi^T Phi)
//
// The reference's analytic gradient (Diff_Loss / The_second_term, CPU Util.cpp:257-304) is
// 3 theta Phi (G o G) - 3 T(phi,phi,I): it equals the exact adjoint when y_A = y_B = y_C and
// mu_A = mu_B = mu_C, and the GPU update direction (stpm.c:649-691) is that expression / 3 with
// theta = 1 in the scaling of init_buff. With A = Y_A Phi, B = Y_B Phi, C = Y_C Phi (b x k):
//   T(I,phi,phi) = Y_A^T K_A / b + mu_A kappa_A^T,  K_A = c3 B o C + c1 (B o m_C + m_B o C),
//                                                   kappa_A = c0 m_B o m_C + c1 mean_x(B o C)
// and likewise for the B and C modes. Everything accumulates in FP64; Y may be FP32 or FP64.
#include <mutex>

#include "common.h"
#include "dense.h"

namespace mmsb {

// per column j: the three K matrices, and the column means needed for kappa and the loss
__global__ void adj_elem_kernel(int32_t nb, int32_t k, const double* __restrict__ A, const double* __restrict__ B,
                                const double* __restrict__ C, const double* __restrict__ m, double c3, double c1,
                                double* __restrict__ KA, double* __restrict__ KB, double* __restrict__ KC,
                                double* __restrict__ colmean) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k) return;
  const double ma = m[j], mb = m[k + j], mc = m[2 * k + j];
  double sbc = 0.0, sac = 0.0, sab = 0.0, st = 0.0;
  for (int x = 0; x < nb; ++x) {
    const int64_t o = (int64_t)x * k + j;
    const double a = A[o], b = B[o], c = C[o];
    KA[o] = c3 * b * c + c1 * (b * mc + mb * c);
    KB[o] = c3 * a * c + c1 * (a * mc + ma * c);
    KC[o] = c3 * a * b + c1 * (a * mb + ma * b);
    sbc += b * c;
    sac += a * c;
    sab += a * b;
    st += c3 * a * b * c + c1 * (a * b * mc + a * mb * c + ma * b * c);
  }
  const double inv = 1.0 / (double)nb;
  colmean[j] = sbc * inv;
  colmean[k + j] = sac * inv;
  colmean[2 * k + j] = sab * inv;
  colmean[3 * k + j] = st * inv;
}

// grad = 3 theta PhiH - T; T gets + mu_X kappa_X^T; loss = theta/2 sum G^3 - sum_i (mean_i + c0 ma mb mc)
__global__ __launch_bounds__(256) void adj_final_kernel(int32_t k, const double* __restrict__ m, const double* __restrict__ mus,
                                                        const double* __restrict__ colmean, const double* __restrict__ G,
                                                        const double* __restrict__ PhiH, const double* __restrict__ T,
                                                        double c1, double c0, double theta, double* __restrict__ grad,
                                                        double* __restrict__ loss) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  const int64_t kk = (int64_t)k * k;
  for (int64_t e = t; e < kk; e += blockDim.x) {
    const int i = (int)(e / k), j = (int)(e % k);   // row i (component), column j (eigenvector)
    const double ma = m[j], mb = m[k + j], mc = m[2 * k + j];
    const double kA = c0 * mb * mc + c1 * colmean[j];
    const double kB = c0 * ma * mc + c1 * colmean[k + j];
    const double kC = c0 * ma * mb + c1 * colmean[2 * k + j];
    const double tt = T[e] + mus[i] * kA + mus[k + i] * kB + mus[2 * k + i] * kC;
    grad[e] = 3.0 * theta * PhiH[e] - tt;
  }
  double acc = 0.0;
  for (int64_t e = t; e < kk; e += blockDim.x) {
    const double g = G[e];
    acc += 0.5 * theta * g * g * g;
  }
  for (int j = t; j < k; j += blockDim.x) acc -= colmean[3 * k + j] + c0 * m[j] * m[k + j] * m[2 * k + j];
  red[t] = acc;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (t < off) red[t] += red[t + off];
    __syncthreads();
  }
  if (t == 0) loss[0] = red[0];
}

__global__ void square_kernel(int64_t n, const double* __restrict__ in, double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = in[t] * in[t];
}

__global__ void pack_mu_kernel(int32_t k, const double* __restrict__ a, const double* __restrict__ b, const double* __restrict__ c,
                               double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < k) {
    out[t] = a[t];
    out[k + t] = b[t];
    out[2 * k + t] = c[t];
  }
}

// Power-method pieces: K_A = c3 B o C + c1 (B o m_C + m_B o C) for a chunk of rows, and the column
// sums of B o C accumulated across chunks.
__global__ void pm_elem_kernel(int32_t nb, int32_t k, const double* __restrict__ B, const double* __restrict__ C,
                               const double* __restrict__ m, double c3, double c1, double* __restrict__ KA,
                               double* __restrict__ colsum) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k) return;
  const double mb = m[k + j], mc = m[2 * k + j];
  double sbc = 0.0;
  for (int x = 0; x < nb; ++x) {
    const int64_t o = (int64_t)x * k + j;
    const double b = B[o], c = C[o];
    KA[o] = c3 * b * c + c1 * (b * mc + mb * c);
    sbc += b * c;
  }
  colsum[j] += sbc;
}

// T[:, j] = T[:, j] / n + mu_A kappa_j, kappa = c0 m_B o m_C + c1 colsum / n;
// lam_j = ||T[:, j]||^3, the reference's (sum phi^2)^(3/2) (Util.cpp:164-165); Phi[:, j] = T[:, j] / ||T[:, j]||
// (normc, Util.cpp:166). One workgroup.
__global__ __launch_bounds__(256) void pm_finish_kernel(int32_t k, double inv_n, const double* __restrict__ T,
                                                        const double* __restrict__ mua, const double* __restrict__ m,
                                                        const double* __restrict__ colsum, double c1, double c0,
                                                        double* __restrict__ phi, double* __restrict__ lam) {
  for (int j = threadIdx.x; j < k; j += blockDim.x) {
    const double kap = c0 * m[k + j] * m[2 * k + j] + c1 * colsum[j] * inv_n;
    double nn = 0.0;
    for (int i = 0; i < k; ++i) {
      const double v = T[(int64_t)i * k + j] * inv_n + mua[i] * kap;
      nn += v * v;
    }
    const double l = sqrt(nn);
    lam[j] = l * l * l;
    const double sc = l > 0.0 ? 1.0 / l : 0.0;
    for (int i = 0; i < k; ++i) phi[(int64_t)i * k + j] = (T[(int64_t)i * k + j] * inv_n + mua[i] * kap) * sc;
  }
}

__global__ void zero_d_kernel(int64_t n, double* p) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.0;
}

// P = U V^T over the singular triplets with sigma_l > 0 (column-major U, V of the one-sided Jacobi
// SVD; out column-major). Applied to the SVD of Phi^T (the row-major Phi buffer read column-major),
// out read row-major is the polar factor of Phi, the orthonormal matrix nearest to it.
__global__ void polar_kernel(int32_t k, const double* __restrict__ U, const double* __restrict__ V,
                             const double* __restrict__ sig, double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)k * k) return;
  const int i = (int)(t % k), j = (int)(t / k);
  double acc = 0.0;
  for (int l = 0; l < k; ++l)
    if (sig[l] > 0.0) acc += U[(int64_t)l * k + i] * V[(int64_t)l * k + j];
  out[t] = acc;
}

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

size_t power_workspace_bytes(int32_t k, int32_t chunk) {
  const size_t ck = sizeof(double) * (size_t)chunk * k, kk = sizeof(double) * (size_t)k * k;
  size_t tn = std::max(tn_workspace_bytes(chunk, k, k), tn_workspace_bytes(k, k, k));
  tn = std::max(tn, tn_workspace_bytes(k, k, 1));
  return 3 * align256(ck) + 4 * align256(kk) + 3 * align256(sizeof(double) * 4 * k) + align256(tn) +
         align256(jacobi_workspace_bytes(k)) + align256(64);
}

// One iteration of the batch implicit power method with the A mode free (CPU tensorDecom_alpha0,
// Util.cpp:131-206): Phi_j <- T(I, phi_j, phi_j) over all n rows, lambda_j = ||.||^3, columns
// normalised, then (optionally) the orthonormalisation Phi <- U V^T from the one-sided Jacobi SVD of
// Phi itself (the polar factor Phi (Phi^T Phi)^-1/2, without forming Phi^T Phi, whose condition
// number is cond(Phi)^2).
// Y_A, Y_B row-major [n x ldy]; Y_C given transposed [k x ldyc] (node-contiguous).
int power_iteration(hipStream_t s, int32_t k, int64_t n, const PowerRows& rows, int32_t ldy, const double* mua,
                    const double* mub, const double* muc, double c3, double c1, double c0, double* phi, double* lam,
                    int orthogonalize, int32_t chunk, void* work) {
  unsigned char* w = static_cast<unsigned char*>(work);
  auto take = [&](size_t bytes) { void* p = w; w += align256(bytes); return p; };
  const size_t ck = sizeof(double) * (size_t)chunk * k, kk = sizeof(double) * (size_t)k * k;
  double *B = (double*)take(ck), *C = (double*)take(ck), *KA = (double*)take(ck);
  double *T = (double*)take(kk), *G = (double*)take(kk), *Vj = (double*)take(kk), *Uj = (double*)take(kk);
  double *m = (double*)take(sizeof(double) * 4 * k), *colsum = (double*)take(sizeof(double) * 4 * k),
         *sig = (double*)take(sizeof(double) * 4 * k);
  double* tn = (double*)take(std::max(tn_workspace_bytes(chunk, k, k), std::max(tn_workspace_bytes(k, k, k), tn_workspace_bytes(k, k, 1))));
  void* jw = take(jacobi_workspace_bytes(k));
  int32_t* sweeps = (int32_t*)take(64);
  MMSB_TRY((gemm_tn<double, double>(s, k, k, 1, 1.0, phi, k, mub, 1, 0.0, m + k, tn)));
  MMSB_TRY((gemm_tn<double, double>(s, k, k, 1, 1.0, phi, k, muc, 1, 0.0, m + 2 * k, tn)));
  hipLaunchKernelGGL(zero_d_kernel, dim3((unsigned)ceil_div(k, 256)), dim3(256), 0, s, (int64_t)k, colsum);
  MMSB_LAUNCH_CHECK();
  for (int64_t r0 = 0; r0 < n; r0 += chunk) {
    const int32_t nb = (int32_t)std::min<int64_t>(chunk, n - r0);
    const float *ya = nullptr, *yb = nullptr, *yc = nullptr;
    MMSB_TRY(rows(r0, nb, &ya, &yb, &yc));
    MMSB_TRY((gemm_nn<float, double>(s, nb, k, k, 1.0, yb, ldy, phi, k, 0.0, B, k)));
    MMSB_TRY((gemm_nn<float, double>(s, nb, k, k, 1.0, yc, ldy, phi, k, 0.0, C, k)));
    hipLaunchKernelGGL(pm_elem_kernel, dim3((unsigned)ceil_div(k, 64)), dim3(64), 0, s, nb, k, B, C, m, c3, c1, KA, colsum);
    MMSB_LAUNCH_CHECK();
    MMSB_TRY((gemm_tn<float, double>(s, nb, k, k, 1.0, ya, ldy, KA, k, r0 == 0 ? 0.0 : 1.0, T, tn)));
  }
  hipLaunchKernelGGL(pm_finish_kernel, dim3(1), dim3(256), 0, s, k, 1.0 / (double)n, T, mua, m, colsum, c1, c0, phi, lam);
  MMSB_LAUNCH_CHECK();
  if (orthogonalize) {  // Phi <- U V^T of Phi's own SVD (G holds a copy: the Jacobi SVD overwrites its input)
    hipError_t e = hipMemcpyAsync(G, phi, kk, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return set_error(MMSB_ERR_HIP, "power: %s", hipGetErrorString(e));
    MMSB_TRY(jacobi_svd(s, k, G, Uj, sig, Vj, jw, sweeps));
    hipLaunchKernelGGL(polar_kernel, dim3((unsigned)ceil_div((int64_t)k * k, 256)), dim3(256), 0, s, k, Uj, Vj, sig, phi);
    MMSB_LAUNCH_CHECK();
  }
  return MMSB_OK;
}

size_t grad_workspace_bytes(int32_t k, int32_t nb) {
  const size_t bk = sizeof(double) * (size_t)nb * k, kk = sizeof(double) * (size_t)k * k;
  size_t tn = std::max(tn_workspace_bytes(nb, k, k), tn_workspace_bytes(k, k, k));
  tn = std::max(tn, tn_workspace_bytes(k, k, 1));
  return 6 * align256(bk) + 4 * align256(kk) + 3 * align256(sizeof(double) * 4 * k) + align256(tn);
}

template <typename TY>
int stgd_grad(hipStream_t s, int32_t k, int32_t nb, const TY* ya, const TY* yb, const TY* yc, int32_t ldy,
              const double* mua, const double* mub, const double* muc, const double* phi, double c3, double c1, double c0,
              double theta, double* grad, double* loss, void* work) {
  MMSB_CHECK_ARG(k >= 1 && nb >= 1 && ldy >= k, "grad: bad shape k=%d nb=%d ldy=%d", k, nb, ldy);
  MMSB_CHECK_ARG(ya && yb && yc && mua && mub && muc && phi && grad && loss && work, "grad: null pointer");
  unsigned char* w = static_cast<unsigned char*>(work);
  auto take = [&](size_t bytes) { double* p = reinterpret_cast<double*>(w); w += align256(bytes); return p; };
  const size_t bk = sizeof(double) * (size_t)nb * k, kk = sizeof(double) * (size_t)k * k;
  double *A = take(bk), *B = take(bk), *C = take(bk), *KA = take(bk), *KB = take(bk), *KC = take(bk);
  double *G = take(kk), *H = take(kk), *PhiH = take(kk), *T = take(kk);
  double *m = take(sizeof(double) * 4 * k), *mus = take(sizeof(double) * 4 * k), *colmean = take(sizeof(double) * 4 * k);
  double* tn = reinterpret_cast<double*>(w);
  // A = Y_A Phi etc. (b x k), m = Phi^T mu
  MMSB_TRY((gemm_nn<TY, double>(s, nb, k, k, 1.0, ya, ldy, phi, k, 0.0, A, k)));
  MMSB_TRY((gemm_nn<TY, double>(s, nb, k, k, 1.0, yb, ldy, phi, k, 0.0, B, k)));
  MMSB_TRY((gemm_nn<TY, double>(s, nb, k, k, 1.0, yc, ldy, phi, k, 0.0, C, k)));
  MMSB_TRY((gemm_tn<double, double>(s, k, k, 1, 1.0, phi, k, mua, 1, 0.0, m, tn)));
  MMSB_TRY((gemm_tn<double, double>(s, k, k, 1, 1.0, phi, k, mub, 1, 0.0, m + k, tn)));
  MMSB_TRY((gemm_tn<double, double>(s, k, k, 1, 1.0, phi, k, muc, 1, 0.0, m + 2 * k, tn)));
  hipLaunchKernelGGL(pack_mu_kernel, dim3((unsigned)ceil_div(k, 256)), dim3(256), 0, s, k, mua, mub, muc, mus);
  MMSB_LAUNCH_CHECK();
  hipLaunchKernelGGL(adj_elem_kernel, dim3((unsigned)ceil_div(k, 64)), dim3(64), 0, s, nb, k, A, B, C, m, c3, c1, KA, KB, KC,
                     colmean);
  MMSB_LAUNCH_CHECK();
  // T = (Y_A^T K_A + Y_B^T K_B + Y_C^T K_C) / b
  const double inv = 1.0 / (double)nb;
  MMSB_TRY((gemm_tn<TY, double>(s, nb, k, k, inv, ya, ldy, KA, k, 0.0, T, tn)));
  MMSB_TRY((gemm_tn<TY, double>(s, nb, k, k, inv, yb, ldy, KB, k, 1.0, T, tn)));
  MMSB_TRY((gemm_tn<TY, double>(s, nb, k, k, inv, yc, ldy, KC, k, 1.0, T, tn)));
  // G = Phi^T Phi, H = G o G, PhiH
  MMSB_TRY((gemm_tn<double, double>(s, k, k, k, 1.0, phi, k, phi, k, 0.0, G, tn)));
  hipLaunchKernelGGL(square_kernel, dim3((unsigned)ceil_div((int64_t)k * k, 256)), dim3(256), 0, s, (int64_t)k * k, G, H);
  MMSB_LAUNCH_CHECK();
  MMSB_TRY((gemm_nn<double, double>(s, k, k, k, 1.0, phi, k, H, k, 0.0, PhiH, k)));
  hipLaunchKernelGGL(adj_final_kernel, dim3(1), dim3(256), 0, s, k, m, mus, colmean, G, PhiH, T, c1, c0, theta, grad, loss);
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

template int stgd_grad<float>(hipStream_t, int32_t, int32_t, const float*, const float*, const float*, int32_t, const double*,
                              const double*, const double*, const double*, double, double, double, double, double*, double*,
                              void*);
template int stgd_grad<double>(hipStream_t, int32_t, int32_t, const double*, const double*, const double*, int32_t,
                               const double*, const double*, const double*, const double*, double, double, double, double,
                               double*, double*, void*);

}  // namespace mmsb

extern "C" size_t mmsb_stgd_grad_workspace(int32_t k, int32_t nb) {
  if (k < 1 || nb < 1) return 0;
  return mmsb::grad_workspace_bytes(k, nb);
}

extern "C" int mmsb_tensor_coeffs(double alpha0, int32_t scaling, double* out3) {
  mmsb::clear_error();
  MMSB_CHECK_ARG(out3 && alpha0 > 0.0, "tensor_coeffs: bad arguments");
  if (scaling == MMSB_M2_GPU) {  // init_buff, stpm.c:1026-1038
    out3[0] = 1.0;
    out3[1] = -alpha0 / (alpha0 + 2.0);
    out3[2] = 2.0 * alpha0 * alpha0 / ((alpha0 + 1.0) * (alpha0 + 2.0));
  } else {  // The_second_term, CPU Util.cpp:289-298
    out3[0] = (alpha0 + 1.0) * (alpha0 + 2.0) / 2.0;
    out3[1] = -alpha0 * (alpha0 + 1.0) / 2.0;
    out3[2] = alpha0 * alpha0;
  }
  return MMSB_OK;
}

extern "C" int mmsb_stgd_grad_f64(void* stream, int32_t k, int32_t nb, const double* ya, const double* yb, const double* yc,
                                  int32_t ldy, const double* mu_a, const double* mu_b, const double* mu_c, const double* phi,
                                  const double* coeffs, double theta, double* grad, double* loss, void* work) {
  mmsb::clear_error();
  MMSB_CHECK_ARG(coeffs, "grad: null coefficients");
  return mmsb::stgd_grad<double>(static_cast<hipStream_t>(stream), k, nb, ya, yb, yc, ldy, mu_a, mu_b, mu_c, phi, coeffs[0],
                                 coeffs[1], coeffs[2], theta, grad, loss, work);
}

namespace mmsb {
// Scratch of the allocation-free adjoint entry point (mmsb_stgd_grad_f32_noalloc): a device array of
// fixed size, one copy per GPU, held until the process ends. Host code never passes the array's name
// to a kernel (on the host it is only a stand-in): the device address of each GPU's copy comes from
// hipGetSymbolAddress, called once per device with that device current, and is kept per device.
static constexpr size_t kGradScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[kGradScratchBytes];
static constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_dev[kMaxScratchDevices] = {};

static int grad_scratch(void** out) {
  int dev = 0;
  MMSB_HIP(hipGetDevice(&dev));
  MMSB_CHECK_ARG(dev >= 0 && dev < kMaxScratchDevices, "grad_noalloc: device %d beyond the supported %d", dev,
                 kMaxScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(MMSB_ERR_HIP, "grad_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return MMSB_OK;
}
}  // namespace mmsb

extern "C" int mmsb_stgd_grad_f32_noalloc(void* stream, int32_t k, int32_t nb, const float* ya, const float* yb,
                                          const float* yc, int32_t ldy, const double* mu_a, const double* mu_b,
                                          const double* mu_c, const double* phi, const double* coeffs, double theta,
                                          double* grad, double* loss) {
  mmsb::clear_error();
  MMSB_CHECK_ARG(coeffs, "grad: null coefficients");
  MMSB_CHECK_ARG(k >= 1 && nb >= 1, "grad_noalloc: bad shape k=%d nb=%d", k, nb);
  const size_t need = mmsb::grad_workspace_bytes(k, nb);
  MMSB_CHECK_ARG(need <= mmsb::kGradScratchBytes,
                 "grad_noalloc: k=%d, nb=%d needs %zu bytes of scratch, above the fixed %zu (use mmsb_stgd_grad_f32 with a "
                 "workspace)", k, nb, need, mmsb::kGradScratchBytes);
  void* work = nullptr;
  MMSB_TRY(mmsb::grad_scratch(&work));
  // every region of the workspace is written before it is read (no clearing needed)
  return mmsb::stgd_grad<float>(static_cast<hipStream_t>(stream), k, nb, ya, yb, yc, ldy, mu_a, mu_b, mu_c, phi, coeffs[0],
                                coeffs[1], coeffs[2], theta, grad, loss, work);
}

extern "C" int mmsb_stgd_grad_f32(void* stream, int32_t k, int32_t nb, const float* ya, const float* yb, const float* yc,
                                  int32_t ldy, const double* mu_a, const double* mu_b, const double* mu_c, const double* phi,
                                  const double* coeffs, double theta, double* grad, double* loss, void* work) {
  mmsb::clear_error();
  MMSB_CHECK_ARG(coeffs, "grad: null coefficients");
  return mmsb::stgd_grad<float>(static_cast<hipStream_t>(stream), k, nb, ya, yb, yc, ldy, mu_a, mu_b, mu_c, phi, coeffs[0],
                                coeffs[1], coeffs[2], theta, grad, loss, work);
}
