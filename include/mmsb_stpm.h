/*
 * mmsb_stpm.h -- C-ABI of the MI355X-native MMSB tensor-method community-detection library.
 *
 * Drop-in boundary for the reference's GPU path (CommunityDetectionSoftware_GPU):
 *   - the compile-time configuration of stpm.h:47-106 becomes the runtime struct stpm_config,
 *     whose field names mirror the macros (NX..NCOM, ALPHA_0, LEARN_RATE, THRESH, PINV_TOL,
 *     SGD_ITER_MIN/MAX, BINARY/EXPECTED, SYNTHETIC, FOLDER_NAME);
 *   - the program-level entry points pre_proc() (pre_proc.c:112), GPU_Task() (stpm.c:132) and
 *     post_proc() (post_proc.c:62) keep their names and read/write the same files;
 *   - the per-kernel launchers of wrappers.cu:148-198 and the CULA/cuRAND call sites
 *     (stpm.c:532-809, gemm_pinv_nys.c:27-421, curand_rand_mat.cu:30-53) are replaced by the
 *     mmsb_* entry points below, which take explicit device pointers and a hipStream_t.
 *
 * Conventions: every function returns an int status (MMSB_OK == 0, negative on error) and never
 * calls exit(); mmsb_last_error() returns the message of the last failure on the calling thread.
 * Device pointers are plain pointers into HBM; `stream` is a hipStream_t passed as void* (NULL is
 * the default stream). Dense tall matrices are row-major (one row per node) with a leading
 * dimension `ld*` in elements; k x k matrices are stated per function.
 */
#ifndef MMSB_STPM_H
#define MMSB_STPM_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (replace exit(1) I/O errors and exit(2) CULA errors, pre_proc.c:54-58,96-109) */
enum {
  MMSB_OK = 0,
  MMSB_ERR_ARG = -1,      /* invalid argument / shape                                   */
  MMSB_ERR_IO = -2,       /* file missing or unreadable (reference: exit(1))            */
  MMSB_ERR_HIP = -3,      /* HIP runtime failure (reference: cula_exception -> exit(2)) */
  MMSB_ERR_NOMEM = -4,    /* host or device allocation failed                           */
  MMSB_ERR_NUMERIC = -5,  /* factorisation failed (not positive definite, no convergence)*/
  MMSB_ERR_STATE = -6     /* entry point called out of order                            */
};

/* Message of the last failing call on this thread ("" if none). */
const char* mmsb_last_error(void);
/* Library version string. */
const char* mmsb_version(void);

/* ------------------------------------------------------------------------------------------
 * Kernel-level entry points
 * ------------------------------------------------------------------------------------------ */

/*
 * CSR SpMM, FP32:  Y[m x r] = alpha * G[m x n] * B[n x r].
 * Replaces every sparse-times-dense product of the whitening, y-vector and Pi stages, which the
 * reference runs as dense n x n DGEMMs through culaBlockDgemm (gemm_pinv_nys.c:27-145; call
 * sites pre_proc.c:154-366, post_proc.c:87-106).
 * rowptr[m+1], colidx[nnz] (int32, 0-based, every colidx < n), vals[nnz] or NULL for a binary
 * graph (all ones). B and Y row-major with leading dimensions ldb, ldy (multiples of 4 floats,
 * 16-byte aligned base pointers), r <= ldb, r <= ldy. Y is overwritten.
 */
int mmsb_spmm_f32(void* stream, int32_t m, int32_t n, int64_t nnz, const int32_t* rowptr,
                  const int32_t* colidx, const float* vals, float alpha, const float* B,
                  int32_t ldb, int32_t r, float* Y, int32_t ldy);

#ifdef __cplusplus
}
#endif

#endif /* MMSB_STPM_H */
