// Pinned double-buffered staging for large host <-> device copies (see xfer.h).
#include "xfer.h"

#include <algorithm>
#include <cstring>
#include <mutex>

#include "host_util.h"
#include "mmsb_stpm.h"

namespace mmsb {

namespace {

constexpr size_t kChunk = (size_t)32 << 20;  // bytes per staging buffer
constexpr size_t kDirect = (size_t)1 << 20;  // below this, plain hipMemcpyAsync

struct Stager {
  std::mutex mu;
  void* buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  int device = -1;
};

Stager& stager() {
  static Stager st;  // kept between copies; freed by release_staging (mmsb_release)
  return st;
}

void drop(Stager& st) {
  for (int i = 0; i < 2; ++i) {
    if (st.ev[i]) (void)hipEventSynchronize(st.ev[i]);
    if (st.buf[i]) (void)hipHostFree(st.buf[i]);
    if (st.ev[i]) (void)hipEventDestroy(st.ev[i]);
    st.buf[i] = nullptr;
    st.ev[i] = nullptr;
  }
  st.device = -1;
}

int ensure(Stager& st) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return set_error(MMSB_ERR_HIP, "staging: hipGetDevice failed");
  if (st.buf[0] && st.device == dev) return MMSB_OK;
  drop(st);
  for (int i = 0; i < 2; ++i) {
    if (hipHostMalloc(&st.buf[i], kChunk, hipHostMallocDefault) != hipSuccess)
      return set_error(MMSB_ERR_NOMEM, "staging: hipHostMalloc(%zu) failed", kChunk);
    if (hipEventCreateWithFlags(&st.ev[i], hipEventDisableTiming) != hipSuccess)
      return set_error(MMSB_ERR_HIP, "staging: hipEventCreate failed");
  }
  st.device = dev;
  return MMSB_OK;
}

void par_copy(void* dst, const void* src, size_t n) {
  const int threads = n >= ((size_t)4 << 20) ? default_threads() : 1;
  parallel_for((int64_t)n, threads, [&](int64_t lo, int64_t hi) {
    std::memcpy(static_cast<char*>(dst) + lo, static_cast<const char*>(src) + lo, (size_t)(hi - lo));
  });
}

}  // namespace

void release_staging() {
  Stager& st = stager();
  std::lock_guard<std::mutex> lock(st.mu);
  drop(st);
}

int h2d(hipStream_t s, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return MMSB_OK;
  if (bytes < kDirect) {
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return set_error(MMSB_ERR_HIP, "h2d: copy failed");
    return MMSB_OK;
  }
  Stager& st = stager();
  std::lock_guard<std::mutex> lock(st.mu);
  if (int e = ensure(st)) return e;
  for (size_t off = 0, i = 0; off < bytes; off += kChunk, ++i) {
    const int b = (int)(i & 1);
    const size_t n = std::min(kChunk, bytes - off);
    if (hipEventSynchronize(st.ev[b]) != hipSuccess) return set_error(MMSB_ERR_HIP, "h2d: event wait failed");
    par_copy(st.buf[b], static_cast<const char*>(src) + off, n);
    if (hipMemcpyAsync(static_cast<char*>(dst) + off, st.buf[b], n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipEventRecord(st.ev[b], s) != hipSuccess)
      return set_error(MMSB_ERR_HIP, "h2d: copy failed");
  }
  return MMSB_OK;
}

int d2h(hipStream_t s, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return MMSB_OK;
  if (bytes < kDirect) {
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
      return set_error(MMSB_ERR_HIP, "d2h: copy failed");
    return MMSB_OK;
  }
  Stager& st = stager();
  std::lock_guard<std::mutex> lock(st.mu);
  if (int e = ensure(st)) return e;
  const size_t nch = (bytes + kChunk - 1) / kChunk;
  auto issue = [&](size_t i) -> int {
    const int b = (int)(i & 1);
    const size_t off = i * kChunk, n = std::min(kChunk, bytes - off);
    if (hipMemcpyAsync(st.buf[b], static_cast<const char*>(src) + off, n, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipEventRecord(st.ev[b], s) != hipSuccess)
      return set_error(MMSB_ERR_HIP, "d2h: copy failed");
    return MMSB_OK;
  };
  if (int e = issue(0)) return e;
  for (size_t i = 0; i < nch; ++i) {
    const int b = (int)(i & 1);
    if (hipEventSynchronize(st.ev[b]) != hipSuccess) return set_error(MMSB_ERR_HIP, "d2h: event wait failed");
    if (i + 1 < nch)
      if (int e = issue(i + 1)) return e;  // the other buffer: its previous chunk was copied out
    const size_t off = i * kChunk, n = std::min(kChunk, bytes - off);
    par_copy(static_cast<char*>(dst) + off, st.buf[b], n);
  }
  return MMSB_OK;
}

}  // namespace mmsb
