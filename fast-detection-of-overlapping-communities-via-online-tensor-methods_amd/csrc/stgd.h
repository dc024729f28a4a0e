// Internal declarations of the STGD kernels (stgd.hip, stgd_bf16.hip).
#pragma once

#include <hip/hip_runtime_api.h>
#include <cstdint>

namespace mmsb {

// Arguments shared by the phase kernels of one STGD step; passed by value to every launch.
struct StgdArgs {
  int32_t kp;           // k padded to a multiple of 64
  int32_t b;            // minibatch tile rows (multiple of 64, >= bl)
  int32_t bl;           // logical minibatch: step s covers rows [s*bl, s*bl + bl) of X
  int64_t n_x;          // nodes in X
  int64_t n_pad;        // rows of Y_A/Y_B and columns of Y_C^T (>= (ceil(n_x / bl) - 1) * bl + b)
  float c3, c1, c0;     // tensor coefficients (GPU scaling of init_buff)
  float eta, theta;     // LEARN_RATE and the weight of the orthogonality term
  const double* mu_a;   // [kp] whitened means
  const double* mu_b;
  const double* mu_c;
  const float* phi;     // [kp x kp] current Phi (row-major, column i = eigenvector i)
  float* phit;          // [kp x kp] Phi~ after the gradient step
  float* h;             // [kp x kp] (Phi^T Phi) o (Phi^T Phi)
  float* cgt;           // [kp x b] C_g transposed
  float* cs_part;       // [b/64 x kp] per-row-tile column sums of P_A o P_B
  float* mab;           // [2 x kp] m_A, m_B
};

// Runs nsteps minibatch steps starting at global step step0 (batch = step mod steps-per-epoch).
// phi_buf[cur] holds Phi on entry; on exit phi_buf[cur] holds the result. frob_log[st] receives
// ||Phi_{t+1} - Phi_t||_F^2 of the st-th step of this call. Every launch is capturable in a HIP graph.
int stgd_run_f32(hipStream_t s, const StgdArgs& args, const float* YA, const float* YB, const float* YCt, int64_t step0,
                 int64_t nsteps, int32_t renorm_every, int32_t ns_iters, float* phi_buf[2], int& cur, float* phit, float* gt,
                 float* phia, double* frob_log);

// BF16 variant (stgd_bf16.hip): Y in BF16; phiT/phitT/phiaT are [kp x kp] FP32 transposes kept beside
// phi_buf/phit/phia; cgt16 is C_g^T in BF16. Requires bl % 8 == 0 (16-byte aligned Y_C^T slices).
// BF16 buffers cross the host boundary as opaque 16-bit words.
typedef uint16_t bf16_bits;
int stgd_run_bf16(hipStream_t s, const StgdArgs& args, const bf16_bits* YA, const bf16_bits* YB, const bf16_bits* YCt,
                  int64_t step0, int64_t nsteps, int32_t renorm_every, int32_t ns_iters, float* phi_buf[2], int& cur,
                  float* phit, float* gt, float* phia, float* phiT, float* phitT, float* phiaT, bf16_bits* cgt16,
                  double* frob_log);
int f32_to_bf16(hipStream_t s, int64_t n, const float* in, bf16_bits* out);

}  // namespace mmsb
