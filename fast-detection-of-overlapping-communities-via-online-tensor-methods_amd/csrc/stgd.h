// Internal declarations of the STGD kernels (stgd.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <functional>
#include <vector>

namespace mmsb {

// Device state of the STGD loop's end test (stpm.c:747-765): steps with a global index >= stop_at are
// skipped; base is the global index of the current run's first step (written before each run).
struct StgdStop {
  int64_t stop_at;
  int64_t base;
};

// Arguments shared by the phase kernels of one STGD step; passed by value to every launch.
struct StgdArgs {
  int32_t k;            // communities
  int32_t kp;           // k padded to a multiple of 64
  int32_t b;            // minibatch tile rows (multiple of 64, >= bl)
  int32_t bl;           // logical minibatch: step s covers rows [s*bl, s*bl + bl) of X
  int64_t n_x;          // nodes in X
  int64_t n_pad;        // rows of Y_A/Y_B and columns of Y_C^T (>= (ceil(n_x / bl) - 1) * bl + b; >= b with y_window)
  float c3, c1, c0;     // tensor coefficients (of the reference path whose M2 scaling whitened Y)
  float eta, theta;     // LEARN_RATE and the weight of the orthogonality term
  const double* mu_a;   // [kp] whitened means
  const double* mu_b;
  const double* mu_c;
  const float* phi;     // [kp x kp] current Phi (row-major, column i = eigenvector i); set per step
  float* phit;          // [kp x kp] Phi~ after the gradient step; set per step
  float* cs_part;       // [b/32 x kp] per-row-tile column sums of P_A o P_B
  float* mab;           // [2 x kp] m_A, m_B
  int32_t y_window;     // 0: Y holds all of X; 1: Y holds the current minibatch only (rows from 0), refilled
                        // before each step by the caller's fill callback (checkpointed STGD)
  int32_t frob_pre;     // 1: log ||Phi~ - Phi||^2 before the renormalisation (the reference's order,
                        // stpm.c:705-743 then the QR at :774-786); 0: ||Phi_next - Phi||^2 after it
  StgdStop* stop;       // null: every step runs; else the end test below, checked per step on the device
  double thresh;        // THRESH: converged when ||dPhi||^2 < thresh and the step index > iter_min
  int64_t iter_min;     // SGD_ITER_MIN (<= 0: the convergence test is off)
  int64_t iter_max;     // SGD_ITER_MAX: the last step is the first with index > iter_max (<= 0: off)
};

// Device buffers of a run (all [kp x kp] FP32 row-major unless noted). The *T buffers hold the
// transposes of their partners so that every MFMA operand is read along contiguous rows. The *16
// buffers ([2 x kp x kp] 16-bit words: the BF16 hi block, then the lo block) hold the split-BF16
// pairs of their partners; BF16 mode only (may be null in FP32 mode).
struct StgdBuffers {
  float* phi[2];        // Phi ping-pong; phi[cur] is the current iterate
  float* phit;          // Phi~ (gradient step) and Newton-Schulz scratch
  float* phia;          // Newton-Schulz scratch
  float* gt;            // X^T X of the current Newton-Schulz iterate
  float* phiT;          // (phi[cur])^T
  float* phitT;         // phit^T
  float* phiaT;         // phia^T
  float* h;             // (Phi^T Phi) o (Phi^T Phi)
  float* q;             // degree-5 Newton-Schulz: the polynomial Q of the iteration's Gram matrix
  uint16_t* phi16[2];
  uint16_t* phit16;
  uint16_t* phia16;
  uint16_t* phiT16;
  uint16_t* phitT16;
  uint16_t* phiaT16;
  uint16_t* gt16;
  uint16_t* h16;
  uint16_t* q16;
  void* cgt;            // [kp x b] C_g^T: FP32, or BF16 in BF16 mode
  unsigned long long* rowsum;  // [2 x kp] row sums of |X^T X| in 2^-32 fixed point (Newton-Schulz scale guard)
  float* mpart;         // [kp/32 x 2 x kp] per-row-block partial sums of mu_A^T Phi, mu_B^T Phi
  double* fpart;        // [(kp/32)^2] per-tile partial sums of the last finished step's log entry
};

// Optional per-phase timing of direct (uncaptured) steps: ev[0..3] are caller-created events
// recorded before the P1 kernel, before P2, before the renormalisation and after it; ms[ph] accumulates
// the elapsed time of phase ph (P1, P2, renormalisation) over the steps of the call. The call then
// waits for each step to finish before launching the next.
struct StgdTiming {
  hipEvent_t ev[4];
  float ms[3];
};

// Runs nsteps minibatch steps starting at global step step0 (batch = step mod steps-per-epoch).
// Every renorm_every-th step renormalises Phi with ns_iters Newton-Schulz iterations of degree
// ns_degree (3: X <- c X (3I - c^2 G)/2; 5: X <- c X (15I - 10 c^2 G + 3 c^4 G^2)/8).
// bufs.phi[cur] holds Phi on entry; cur is flipped once per step and bufs.phi[cur] holds the
// result on exit. frob_log[st] receives ||Phi_{t+1} - Phi_t||_F^2 of the st-th step of this call.
// precision is MMSB_PREC_FP32 (Y as float) or MMSB_PREC_BF16 (Y as BF16). Without timing every
// launch is capturable in a HIP graph (no memset nodes, no host synchronisation).
// With args.y_window, fill(row0, nb) is called before each step's first launch to enqueue (on s) the
// kernels that write rows [row0, row0 + nb) of X into the Y buffers.
// With a renorm callback (the reference's QR mode), a renormalising step runs P1, P2, then
// renorm(src = bufs.phit, dst = bufs.phia) -- launches that leave the orthonormal factor of Phi~ in
// phia -- and a QRFIN kernel that makes it the next Phi (with its derived copies).
using StgdFill = std::function<int(int64_t row0, int32_t nb)>;
using StgdRenorm = std::function<int(const float* src, float* dst)>;
int stgd_run(hipStream_t s, int32_t precision, const StgdArgs& args, const void* YA, const void* YB, const void* YCt,
             int64_t step0, int64_t nsteps, int32_t renorm_every, int32_t ns_iters, int32_t ns_degree,
             const StgdBuffers& bufs, int& cur, double* frob_log, StgdTiming* timing = nullptr,
             const StgdFill& fill = StgdFill(), const StgdRenorm& renorm = StgdRenorm());
// Sets *p = v on stream s (one thread; capturable, unlike a memset node).
int stgd_set_i64(hipStream_t s, int64_t* p, int64_t v);

// ---- many instances in one persistent launch ----------------------------------------------------
// One member of a batch: its step arguments, buffers and neighbour vectors (FP32 or BF16 as the
// batch precision), the state of its Phi ping-pong, and the steps of one launch.
struct StgdBatchMember {
  StgdArgs args;
  StgdBuffers bufs;
  const void* YA;
  const void* YB;
  const void* YCt;
  int32_t cur;        // bufs.phi[cur] holds Phi at the start of every launch of the plan
  int64_t step0;      // global step index of the launch's first step
  int64_t nsteps;     // steps of one launch (e.g. one epoch)
  double* frob_log;   // [nsteps]
};

struct StgdBatch;
// Lays out the tiles of every member's steps slot by slot and uploads the plan. The plan is fixed:
// every launch replays the same steps (the members' cur must match at each launch).
int stgd_batch_plan(StgdBatch*& out, int32_t precision, const std::vector<StgdBatchMember>& members, int32_t ns_iters,
                    int32_t ns_degree, int32_t renorm_every);
// One cooperative launch of the persistent kernel over the whole plan.
int stgd_batch_launch(hipStream_t s, StgdBatch* b);
// The same plan as one grouped launch per phase slot (capturable in a HIP graph).
int stgd_batch_launch_slots(hipStream_t s, StgdBatch* b);
// Workgroups of the launch (at most every workgroup that can be resident at once).
void stgd_batch_set_grid(StgdBatch* b, int32_t grid);
int32_t stgd_batch_grid(const StgdBatch* b);
int32_t stgd_batch_grid_max(const StgdBatch* b);
void stgd_batch_free(StgdBatch* b);

// BF16 buffers cross the host boundary as opaque 16-bit words.
typedef uint16_t bf16_bits;
int f32_to_bf16(hipStream_t s, int64_t n, const float* in, bf16_bits* out);

}  // namespace mmsb
