// Post-processing on the GPU: eigenvalues of the recovered components, the inverse-eigenvalue
// scaling used to unwhiten, and the per-node membership normalisation and threshold.
//
//   lambda_i = ||phi_i||^3                        (stpm.c:799 gemv of phi_i with itself, :809 pow 1.5)
//   Lambda'_ii = ||lambda||_2 / lambda_i, 0 if |lambda_i| <= PINV_TOL   (genInvEvalMat, wrappers.cu:74-96;
//                                                                      l2Norm, wrappers.cu:121-132)
//   alpha_hat_i = lambda_i^-2 / sum_j lambda_j^-2 (stpm.c:848-870)
//   Pi_hat(node) = normProbVector(Pi(node))        (CPU Probability.cpp:33-57), memberships = {c : Pi_hat > tau}
#include "common.h"
#include "post.h"

namespace mmsb {

// One workgroup: lambda, Lambda', alpha_hat, and PhiL[m][i] = Phi[m][i] * Lambda'_i (FP64, k x k).
__global__ __launch_bounds__(1024) void eig_post_kernel(int32_t k, int32_t kp, const float* __restrict__ phi, double pinv_tol,
                                                        double* __restrict__ lam, double* __restrict__ lamp,
                                                        double* __restrict__ alpha_hat, double* __restrict__ phil) {
  __shared__ double red[2][1024];
  const int t = threadIdx.x;
  double s2 = 0.0, si = 0.0;
  for (int i = t; i < k; i += blockDim.x) {
    double nn = 0.0;
    for (int m = 0; m < k; ++m) {
      const double v = (double)phi[(int64_t)m * kp + i];
      nn = fma(v, v, nn);
    }
    const double l = nn * sqrt(nn);  // (||phi||^2)^1.5
    lam[i] = l;
    s2 += l * l;
    si += 1.0 / (l * l);
  }
  red[0][t] = s2;
  red[1][t] = si;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (t < off) {
      red[0][t] += red[0][t + off];
      red[1][t] += red[1][t + off];
    }
    __syncthreads();
  }
  const double nrm = sqrt(red[0][0]);
  const double sum_inv2 = red[1][0];
  for (int i = t; i < k; i += blockDim.x) {
    const double l = lam[i];
    lamp[i] = fabs(l) > pinv_tol ? nrm / l : 0.0;
    alpha_hat[i] = (1.0 / (l * l)) / sum_inv2;
  }
  __syncthreads();
  for (int64_t e = t; e < (int64_t)k * k; e += blockDim.x) {
    const int m = (int)(e / k), i = (int)(e % k);
    phil[e] = (double)phi[(int64_t)m * kp + i] * lamp[i];
  }
}

int eig_post(hipStream_t s, int32_t k, int32_t kp, const float* phi, double pinv_tol, double* lam, double* lamp,
             double* alpha_hat, double* phil) {
  hipLaunchKernelGGL(eig_post_kernel, dim3(1), dim3(1024), 0, s, k, kp, phi, pinv_tol, lam, lamp, alpha_hat, phil);
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

// One wave per node: normProbVector of the node's k values, clamp, count entries above tau.
__global__ __launch_bounds__(256) void norm_prob_kernel(int64_t n, int32_t k, const float* __restrict__ P, int32_t ldp,
                                                        float* __restrict__ out, int32_t ldo, float tau,
                                                        int32_t* __restrict__ counts) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (row >= n) return;
  const float* p = P + row * ldp;
  double pos = 0.0, neg = 0.0;
  int nz = 0;
  for (int c = lane; c < k; c += 64) {
    const float v = p[c];
    if (v > 0.f) pos += v; else neg += v;
    nz |= (v != 0.f);
  }
  for (int off = 32; off > 0; off >>= 1) {
    pos += __shfl_xor(pos, off);
    neg += __shfl_xor(neg, off);
    nz |= __shfl_xor(nz, off);
  }
  const bool flip = fabs(pos) < fabs(neg);
  double den = flip ? fabs(neg) : fabs(pos);
  if (!nz || den == 0.0) den = 1.0;
  int cnt = 0;
  float* o = out + row * ldo;
  for (int c = lane; c < k; c += 64) {
    double v = (double)p[c];
    v = (flip ? -v : v) / den;
    if (v < 0.0 || !nz) v = 0.0;
    const float f = (float)v;
    o[c] = f;
    cnt += (f > tau) ? 1 : 0;
  }
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
  if (lane == 0 && counts) counts[row] = cnt;
}

int norm_prob(hipStream_t s, int64_t n, int32_t k, const float* P, int32_t ldp, float* out, int32_t ldo, float tau,
              int32_t* counts) {
  if (n <= 0) return MMSB_OK;
  hipLaunchKernelGGL(norm_prob_kernel, dim3((unsigned)ceil_div(n, 4)), dim3(256), 0, s, n, k, P, ldp, out, ldo, tau, counts);
  MMSB_LAUNCH_CHECK();
  return MMSB_OK;
}

}  // namespace mmsb
