// Internal declarations of the post-processing kernels (post.hip).
#pragma once

#include <hip/hip_runtime_api.h>
#include <cstdint>

namespace mmsb {

int eig_post(hipStream_t s, int32_t k, int32_t kp, const float* phi, double pinv_tol, double* lam, double* lamp,
             double* alpha_hat, double* phil);
int norm_prob(hipStream_t s, int64_t n, int32_t k, const float* P, int32_t ldp, float* out, int32_t ldo, float tau,
              int32_t* counts);

}  // namespace mmsb
