// CSR transpose on the device (G^T of each adjacency block, built once after upload), and the
// rows of a CSR matrix ordered by descending length (the SpMM's row schedule).
//
// A stable LSD radix sort (hipCUB) of (column, nonzero index) pairs orders the nonzeros by column
// and, within a column, by their position in G -- that is, by row, because G's nonzeros are
// stored row by row. Row pointers of G^T are the lower bounds of each column in the sorted keys;
// the rows themselves are recovered from the nonzero index by a binary search over G's row
// pointers, and values (if any) gathered by the same index. The result equals the host
// counting-sort transpose element for element.
#include <hipcub/hipcub.hpp>

#include <mutex>

#include "common.h"
#include "dense.h"
#include "xfer.h"

namespace mmsb {

namespace {
// Temporaries of the sorts (keys, indices, the radix sort's own storage), kept between calls and
// grown on demand for the current device: a set of blocks is transposed and ordered without a
// hipMalloc / hipFree pair per temporary. Calls are serialised, and each waits for its stream
// before returning, so the next call finds them idle.
struct SortScratch {
  std::mutex mu;
  int device = -1;
  void* p[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t bytes[4] = {0, 0, 0, 0};
};

SortScratch& sort_scratch() {
  static SortScratch s;  // kept between calls; freed by release_sort_scratch (mmsb_release)
  return s;
}

void drop(SortScratch& s) {
  for (int i = 0; i < 4; ++i) {
    if (s.p[i]) (void)hipFree(s.p[i]);
    s.p[i] = nullptr;
    s.bytes[i] = 0;
  }
  s.device = -1;
}

hipError_t scratch_ready(SortScratch& s) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev == s.device) return e;
  drop(s);
  s.device = dev;
  return hipSuccess;
}

template <typename T>
hipError_t scratch_take(SortScratch& s, int i, size_t bytes, T** out) {
  bytes = std::max<size_t>(bytes, 256);
  if (s.bytes[i] < bytes) {
    if (s.p[i]) (void)hipFree(s.p[i]);
    s.p[i] = nullptr;
    s.bytes[i] = 0;
    const hipError_t e = hipMalloc(&s.p[i], bytes);
    if (e != hipSuccess) return e;
    s.bytes[i] = bytes;
  }
  *out = static_cast<T*>(s.p[i]);
  return hipSuccess;
}
}  // namespace

void release_sort_scratch() {
  SortScratch& s = sort_scratch();
  std::lock_guard<std::mutex> lock(s.mu);
  drop(s);
}

__global__ void iota_kernel(int64_t n, int32_t* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = (int32_t)t;
}

// rpT[c] = number of sorted keys < c, for c in [0, nc]
__global__ void lower_bounds_kernel(int32_t nc, int64_t nnz, const int32_t* __restrict__ keys, int32_t* __restrict__ rpT) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c > nc) return;
  int64_t lo = 0, hi = nnz;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  rpT[c] = (int32_t)lo;
}

// ciT[p] = row of nonzero perm[p] (last row r with rp[r] <= perm[p]); vT[p] = v[perm[p]]
__global__ void gather_rows_kernel(int32_t nr, int64_t nnz, const int32_t* __restrict__ rp, const int32_t* __restrict__ perm,
                                   const float* __restrict__ v, int32_t* __restrict__ ciT, float* __restrict__ vT) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nnz) return;
  const int32_t e = perm[p];
  int32_t lo = 0, hi = nr - 1;  // rp[lo] <= e < rp[nr]
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (rp[mid] <= e) lo = mid;
    else hi = mid - 1;
  }
  ciT[p] = lo;
  if (v) vT[p] = v[e];
}

static int bits_for(int32_t n) {
  int b = 1;
  while (b < 31 && (1ll << b) < (int64_t)n) ++b;
  return b;
}

int csr_transpose(hipStream_t s, int32_t nr, int32_t nc, int64_t nnz, const int32_t* rp, const int32_t* ci, const float* v,
                  int32_t* rpT, int32_t* ciT, float* vT) {
  MMSB_CHECK_ARG(nr >= 1 && nc >= 1 && nnz >= 0 && nnz < INT32_MAX, "csr_transpose: bad shape");
  if (nnz == 0) {
    hipLaunchKernelGGL(lower_bounds_kernel, dim3((unsigned)ceil_div((int64_t)nc + 1, 256)), dim3(256), 0, s, nc, (int64_t)0,
                       ci, rpT);
    MMSB_LAUNCH_CHECK();
    return MMSB_OK;
  }
  int32_t *keys_out = nullptr, *idx = nullptr, *perm = nullptr;
  void* temp = nullptr;
  size_t temp_bytes = 0;
  const int end_bit = bits_for(nc);
  int st = MMSB_OK;
  auto fail = [&](hipError_t e) { return set_error(MMSB_ERR_HIP, "csr_transpose: %s", hipGetErrorString(e)); };
  SortScratch& sc = sort_scratch();
  std::lock_guard<std::mutex> lock(sc.mu);
  hipError_t e = scratch_ready(sc);
  if (e == hipSuccess) e = scratch_take(sc, 0, sizeof(int32_t) * nnz, &keys_out);
  if (e == hipSuccess) e = scratch_take(sc, 1, sizeof(int32_t) * nnz, &idx);
  if (e == hipSuccess) e = scratch_take(sc, 2, sizeof(int32_t) * nnz, &perm);
  if (e == hipSuccess)
    e = hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, ci, keys_out, idx, perm, (int)nnz, 0, end_bit, s);
  if (e == hipSuccess) e = scratch_take(sc, 3, temp_bytes, &temp);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(iota_kernel, dim3((unsigned)ceil_div(nnz, 256)), dim3(256), 0, s, nnz, idx);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, ci, keys_out, idx, perm, (int)nnz, 0, end_bit, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(lower_bounds_kernel, dim3((unsigned)ceil_div((int64_t)nc + 1, 256)), dim3(256), 0, s, nc, nnz,
                       keys_out, rpT);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)ceil_div(nnz, 256)), dim3(256), 0, s, nr, nnz, rp, perm, v, ciT, vT);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);  // the temporaries are idle for the next call
  if (e != hipSuccess) st = fail(e);
  return st;
}

__global__ void row_lengths_kernel(int32_t nr, const int32_t* __restrict__ rp, int32_t* __restrict__ len) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nr) len[i] = rp[i + 1] - rp[i];
}

int csr_row_order(hipStream_t s, int32_t nr, const int32_t* rp, int32_t* order) {
  MMSB_CHECK_ARG(nr >= 1 && rp && order, "csr_row_order: bad arguments");
  int32_t *len = nullptr, *len_out = nullptr, *idx = nullptr;
  void* temp = nullptr;
  size_t temp_bytes = 0;
  int st = MMSB_OK;
  SortScratch& sc = sort_scratch();
  std::lock_guard<std::mutex> lock(sc.mu);
  hipError_t e = scratch_ready(sc);
  if (e == hipSuccess) e = scratch_take(sc, 0, sizeof(int32_t) * nr, &len);
  if (e == hipSuccess) e = scratch_take(sc, 1, sizeof(int32_t) * nr, &len_out);
  if (e == hipSuccess) e = scratch_take(sc, 2, sizeof(int32_t) * nr, &idx);
  if (e == hipSuccess)
    e = hipcub::DeviceRadixSort::SortPairsDescending(nullptr, temp_bytes, len, len_out, idx, order, (int)nr, 0, 31, s);
  if (e == hipSuccess) e = scratch_take(sc, 3, temp_bytes, &temp);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(row_lengths_kernel, dim3((unsigned)ceil_div((int64_t)nr, 256)), dim3(256), 0, s, nr, rp, len);
    hipLaunchKernelGGL(iota_kernel, dim3((unsigned)ceil_div((int64_t)nr, 256)), dim3(256), 0, s, (int64_t)nr, idx);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipcub::DeviceRadixSort::SortPairsDescending(temp, temp_bytes, len, len_out, idx, order, (int)nr, 0, 31, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);  // the temporaries are idle for the next call
  if (e != hipSuccess) st = set_error(MMSB_ERR_HIP, "csr_row_order: %s", hipGetErrorString(e));
  return st;
}

}  // namespace mmsb

extern "C" int mmsb_csr_row_order(void* stream, int32_t m, const int32_t* rowptr, int32_t* order) {
  mmsb::clear_error();
  return mmsb::csr_row_order(static_cast<hipStream_t>(stream), m, rowptr, order);
}
