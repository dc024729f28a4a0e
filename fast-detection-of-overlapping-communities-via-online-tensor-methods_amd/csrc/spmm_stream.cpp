// CSR x dense SpMM with the sparse block in host memory, streamed to the device in row chunks.
//
// The reference moves operand tiles between host and device for every product (culaBlockDgemm,
// gemm_pinv_nys.c:27-145, BLOCK_SIZE tiles of dense n x n buffers). Here only the CSR block stays in
// host memory -- for blocks that are not kept in HBM -- while B and Y are device-resident: the rows
// are cut into chunks of at most `chunk_nnz` nonzeros, each chunk is rebased and staged in pinned
// memory by the host, copied on a copy stream, and multiplied by the resident SpMM kernel on the
// caller's stream. Two slots alternate, so the copy of chunk i + 1 overlaps the product of chunk i.
// When colidx (and vals) are already pinned host memory (hipHostMalloc), their chunks are copied
// straight from there and only the rebased row pointers and the row order go through the staging, so
// a block kept pinned between calls streams at the PCIe rate.
// The host also orders each chunk's rows longest first (a counting sort of the row lengths), the
// schedule the resident blocks use. Each row is summed by the same kernel as with a device-resident
// block, so Y is identical.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <utility>
#include <vector>

#include "dense.h"
#include "host_util.h"
#include "mmsb_stpm.h"
#include "xfer.h"

namespace mmsb {

#define SHIP(expr)                                                                                     \
  do {                                                                                                 \
    hipError_t e_ = (expr);                                                                            \
    if (e_ != hipSuccess)                                                                              \
      return ::mmsb::set_error(MMSB_ERR_HIP, "%s:%d %s -> %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_)); \
  } while (0)
#define STRY(expr)                                                                                     \
  do {                                                                                                 \
    int s_ = (expr);                                                                                   \
    if (s_ != MMSB_OK) return s_;                                                                      \
  } while (0)

namespace {

constexpr int64_t kChunkRows = (int64_t)1 << 20;  // rows per chunk at most (bounds the rowptr slot)

struct Slot {
  void* host = nullptr;         // pinned staging: rowptr, row order, colidx, then vals
  size_t host_bytes = 0;
  int32_t* rowptr = nullptr;    // device
  int32_t* order = nullptr;
  int32_t* colidx = nullptr;
  float* vals = nullptr;
  int64_t rows_cap = 0, nnz_cap = 0;
  hipEvent_t copied = nullptr;  // recorded on the copy stream after the slot's upload
  hipEvent_t used = nullptr;    // recorded on the compute stream after the slot's product
};

// Process-wide staging (grown on demand, kept for later calls); calls are serialised.
struct Streamer {
  std::mutex mu;
  int device = -1;
  hipStream_t copy = nullptr;
  Slot slot[2];
};

Streamer& streamer() {
  static Streamer st;  // kept between calls; freed by release_streamer (mmsb_release)
  return st;
}

// Waits for both slots and frees everything the streamer holds (caller holds st.mu).
void drop(Streamer& st) {
  for (Slot& sl : st.slot) {
    if (sl.used) (void)hipEventSynchronize(sl.used);
    if (sl.copied) (void)hipEventSynchronize(sl.copied);
    for (void* p : {(void*)sl.rowptr, (void*)sl.order, (void*)sl.colidx, (void*)sl.vals})
      if (p) (void)hipFree(p);
    if (sl.host) (void)hipHostFree(sl.host);
    if (sl.copied) (void)hipEventDestroy(sl.copied);
    if (sl.used) (void)hipEventDestroy(sl.used);
    sl = Slot();
  }
  if (st.copy) (void)hipStreamDestroy(st.copy);
  st.copy = nullptr;
  st.device = -1;
}

int ensure_slot(Slot& sl, int64_t rows, int64_t nnz, bool vals) {
  if (!sl.copied) {
    SHIP(hipEventCreateWithFlags(&sl.copied, hipEventDisableTiming));
    SHIP(hipEventCreateWithFlags(&sl.used, hipEventDisableTiming));
    SHIP(hipEventRecord(sl.used, nullptr));  // an initial "free" mark
  }
  if (rows > sl.rows_cap || nnz > sl.nnz_cap || (vals && !sl.vals)) {
    SHIP(hipEventSynchronize(sl.used));
    for (void* p : {(void*)sl.rowptr, (void*)sl.order, (void*)sl.colidx, (void*)sl.vals})
      if (p) (void)hipFree(p);
    sl.rowptr = nullptr;
    sl.order = nullptr;
    sl.colidx = nullptr;
    sl.vals = nullptr;
    sl.rows_cap = std::max(rows, sl.rows_cap);
    sl.nnz_cap = std::max(nnz, sl.nnz_cap);
    SHIP(hipMalloc(&sl.rowptr, sizeof(int32_t) * (size_t)(sl.rows_cap + 1)));
    SHIP(hipMalloc(&sl.order, sizeof(int32_t) * (size_t)sl.rows_cap));
    SHIP(hipMalloc(&sl.colidx, sizeof(int32_t) * (size_t)std::max<int64_t>(sl.nnz_cap, 1)));
    if (vals) SHIP(hipMalloc(&sl.vals, sizeof(float) * (size_t)std::max<int64_t>(sl.nnz_cap, 1)));
  }
  const size_t hb = sizeof(int32_t) * (size_t)(2 * sl.rows_cap + 1) + sizeof(int32_t) * (size_t)sl.nnz_cap +
                    (sl.vals ? sizeof(float) * (size_t)sl.nnz_cap : 0);
  if (hb > sl.host_bytes) {
    SHIP(hipEventSynchronize(sl.copied));
    if (sl.host) (void)hipHostFree(sl.host);
    sl.host = nullptr;
    if (hipHostMalloc(&sl.host, hb, hipHostMallocDefault) != hipSuccess)
      return set_error(MMSB_ERR_NOMEM, "spmm_host_csr: hipHostMalloc(%zu) failed", hb);
    sl.host_bytes = hb;
  }
  return MMSB_OK;
}

// Pinned (page-locked, device-visible) host memory? A pageable pointer makes the query fail; the
// failure is cleared so that it does not surface at a later launch check.
bool is_pinned(const void* p) {
  if (!p) return true;
  hipPointerAttribute_t a{};
  const hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeHost;
}

void par_copy(void* dst, const void* src, size_t n) {
  const int threads = n >= ((size_t)4 << 20) ? default_threads() : 1;
  parallel_for((int64_t)n, threads, [&](int64_t lo, int64_t hi) {
    std::memcpy(static_cast<char*>(dst) + lo, static_cast<const char*>(src) + lo, (size_t)(hi - lo));
  });
}

}  // namespace

int spmm_f32_host_csr(hipStream_t s, int32_t m, int32_t n, int64_t nnz, const int32_t* rowptr, const int32_t* colidx,
                      const float* vals, float alpha, const float* B, int32_t ldb, int32_t r, float* Y, int32_t ldy,
                      int64_t chunk_nnz, bool check) {
  MMSB_CHECK_ARG(m >= 0 && n >= 0 && nnz >= 0 && nnz < (int64_t)INT32_MAX, "spmm_host_csr: bad sizes m=%d n=%d nnz=%lld", m,
                 n, (long long)nnz);
  MMSB_CHECK_ARG(chunk_nnz >= 1, "spmm_host_csr: chunk_nnz must be positive");
  if (m == 0) return MMSB_OK;
  MMSB_CHECK_ARG(rowptr && (nnz == 0 || colidx), "spmm_host_csr: null CSR");
  MMSB_CHECK_ARG(rowptr[0] == 0 && rowptr[m] == nnz, "spmm_host_csr: rowptr does not span [0, nnz]");
  if (check) {  // the block is on the host: check it before any of it reaches the kernel
    const int nt = default_threads();
    std::vector<int64_t> bad_row(nt, -1), bad_nz(nt, -1);
    parallel_chunks(m, nt, [&](int t, int64_t lo, int64_t hi) {
      for (int64_t i = lo; i < hi && bad_row[t] < 0; ++i)
        if (rowptr[i + 1] < rowptr[i]) bad_row[t] = i;
    });
    parallel_chunks(nnz, nt, [&](int t, int64_t lo, int64_t hi) {
      for (int64_t e = lo; e < hi && bad_nz[t] < 0; ++e)
        if (colidx[e] < 0 || colidx[e] >= n) bad_nz[t] = e;
    });
    for (int t = 0; t < nt; ++t) {
      MMSB_CHECK_ARG(bad_row[t] < 0, "spmm_host_csr: rowptr decreases at row %lld", (long long)bad_row[t]);
      MMSB_CHECK_ARG(bad_nz[t] < 0, "spmm_host_csr: column index out of range at nonzero %lld", (long long)bad_nz[t]);
    }
  }
  // chunks of whole rows: at most chunk_nnz nonzeros (a longer row is a chunk of its own) and kChunkRows rows
  std::vector<std::pair<int32_t, int32_t>> chunks;
  int64_t max_rows = 0, max_nnz = 0;
  for (int32_t r0 = 0; r0 < m;) {
    int32_t r1 = r0 + 1;
    while (r1 < m && r1 - r0 < kChunkRows && (int64_t)rowptr[r1 + 1] - rowptr[r0] <= chunk_nnz) ++r1;
    chunks.emplace_back(r0, r1);
    max_rows = std::max<int64_t>(max_rows, r1 - r0);
    max_nnz = std::max<int64_t>(max_nnz, (int64_t)rowptr[r1] - rowptr[r0]);
    r0 = r1;
  }
  Streamer& st = streamer();
  std::lock_guard<std::mutex> lock(st.mu);
  int dev = 0;
  SHIP(hipGetDevice(&dev));
  if (st.device != dev) {  // one set of staging per process, for the current device
    drop(st);
    SHIP(hipStreamCreateWithFlags(&st.copy, hipStreamNonBlocking));
    st.device = dev;
  }
  for (Slot& sl : st.slot) STRY(ensure_slot(sl, max_rows, max_nnz, vals != nullptr));
  const bool direct = is_pinned(colidx) && is_pinned(vals);
  for (size_t i = 0; i < chunks.size(); ++i) {
    Slot& sl = st.slot[i & 1];
    const int32_t r0 = chunks[i].first, r1 = chunks[i].second, rows = r1 - r0;
    const int64_t z0 = rowptr[r0], cnnz = (int64_t)rowptr[r1] - z0;
    // the slot's previous upload has left the staging, its previous product has read the device copy
    SHIP(hipEventSynchronize(sl.copied));
    int32_t* hrp = static_cast<int32_t*>(sl.host);
    int32_t* hord = hrp + (sl.rows_cap + 1);
    int32_t* hci = hord + sl.rows_cap;
    float* hv = reinterpret_cast<float*>(hci + sl.nnz_cap);
    int32_t maxlen = 0;
    for (int32_t j = 0; j <= rows; ++j) hrp[j] = (int32_t)((int64_t)rowptr[r0 + j] - z0);
    for (int32_t j = 0; j < rows; ++j) maxlen = std::max(maxlen, hrp[j + 1] - hrp[j]);
    {  // rows longest first, ties by index: a stable counting sort on the lengths
      std::vector<int32_t> start((size_t)maxlen + 2, 0);
      for (int32_t j = 0; j < rows; ++j) ++start[(size_t)(maxlen - (hrp[j + 1] - hrp[j])) + 1];
      for (size_t b = 1; b < start.size(); ++b) start[b] += start[b - 1];
      for (int32_t j = 0; j < rows; ++j) hord[start[(size_t)(maxlen - (hrp[j + 1] - hrp[j]))]++] = j;
    }
    const int32_t* src_ci = colidx + z0;
    const float* src_v = vals ? vals + z0 : nullptr;
    if (!direct) {
      par_copy(hci, src_ci, sizeof(int32_t) * (size_t)cnnz);
      if (vals) par_copy(hv, src_v, sizeof(float) * (size_t)cnnz);
      src_ci = hci;
      src_v = hv;
    }
    SHIP(hipStreamWaitEvent(st.copy, sl.used, 0));
    SHIP(hipMemcpyAsync(sl.rowptr, hrp, sizeof(int32_t) * (size_t)(rows + 1), hipMemcpyHostToDevice, st.copy));
    SHIP(hipMemcpyAsync(sl.order, hord, sizeof(int32_t) * (size_t)rows, hipMemcpyHostToDevice, st.copy));
    if (cnnz) SHIP(hipMemcpyAsync(sl.colidx, src_ci, sizeof(int32_t) * (size_t)cnnz, hipMemcpyHostToDevice, st.copy));
    if (vals && cnnz) SHIP(hipMemcpyAsync(sl.vals, src_v, sizeof(float) * (size_t)cnnz, hipMemcpyHostToDevice, st.copy));
    SHIP(hipEventRecord(sl.copied, st.copy));
    SHIP(hipStreamWaitEvent(s, sl.copied, 0));
    STRY(spmm_f32(s, rows, n, cnnz, sl.rowptr, sl.colidx, vals ? sl.vals : nullptr, alpha, B, ldb, r,
                      Y + (int64_t)r0 * ldy, ldy, sl.order));
    SHIP(hipEventRecord(sl.used, s));
  }
  SHIP(hipStreamSynchronize(s));
  return MMSB_OK;
}

void release_streamer() {
  Streamer& st = streamer();
  std::lock_guard<std::mutex> lock(st.mu);
  drop(st);
}

}  // namespace mmsb

extern "C" int mmsb_spmm_f32_host_csr(void* stream, int32_t m, int32_t n, int64_t nnz, const int32_t* rowptr,
                                      const int32_t* colidx, const float* vals, float alpha, const float* B, int32_t ldb,
                                      int32_t r, float* Y, int32_t ldy, int64_t chunk_nnz) {
  mmsb::clear_error();
  return mmsb::spmm_f32_host_csr(static_cast<hipStream_t>(stream), m, n, nnz, rowptr, colidx, vals, alpha, B, ldb, r, Y,
                                 ldy, chunk_nnz);
}
