// Host-side helpers: error recording, a static-range parallel_for over std::thread, and a
// bucketed parallel sort + unique for 64-bit keys (deterministic output: the sorted unique set).
#pragma once

#include <algorithm>
#include <cstdint>
#include <functional>
#include <thread>
#include <vector>

#include "mmsb_stpm.h"

namespace mmsb {

int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

#ifndef MMSB_CHECK_ARG
#define MMSB_CHECK_ARG(cond, ...)                                                       \
  do {                                                                                  \
    if (!(cond)) return ::mmsb::set_error(MMSB_ERR_ARG, __VA_ARGS__);                   \
  } while (0)
#endif

inline int default_threads() {
  unsigned h = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(h, 16u));
}

// Calls fn(lo, hi) on `threads` contiguous sub-ranges of [0, n).
// Runs fn(chunk, lo, hi) for the min(threads, n) static chunks [n c / t, n (c + 1) / t) of [0, n),
// one thread each. Per-thread state must be indexed by the chunk argument: recovering it from lo
// (as lo * t / n) is off by one whenever n * c / t is not an integer.
inline void parallel_chunks(int64_t n, int threads, const std::function<void(int, int64_t, int64_t)>& fn) {
  if (n <= 0) return;
  const int64_t t = std::max<int64_t>(1, std::min<int64_t>(threads, n));
  if (t == 1) {
    fn(0, 0, n);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(t);
  for (int64_t i = 0; i < t; ++i) pool.emplace_back(fn, (int)i, n * i / t, n * (i + 1) / t);
  for (auto& th : pool) th.join();
}

inline void parallel_for(int64_t n, int threads, const std::function<void(int64_t, int64_t)>& fn) {
  parallel_chunks(n, threads, [&](int, int64_t lo, int64_t hi) { fn(lo, hi); });
}

// Sorts keys ascending and removes duplicates.
inline void parallel_sort_unique(std::vector<uint64_t>& keys, int threads) {
  const int64_t n = (int64_t)keys.size();
  if (n < (1 << 20) || threads <= 1) {
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    return;
  }
  const uint64_t mx = *std::max_element(keys.begin(), keys.end());
  const int nb = 8 * threads;
  const uint64_t width = mx / (uint64_t)nb + 1;
  std::vector<std::vector<int64_t>> hist(threads, std::vector<int64_t>(nb, 0));
  parallel_chunks(n, threads, [&](int t, int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) hist[t][keys[i] / width]++;
  });
  // offsets per (bucket, thread) in bucket-major order
  std::vector<std::vector<int64_t>> off(threads, std::vector<int64_t>(nb, 0));
  std::vector<int64_t> bstart(nb + 1, 0);
  int64_t pos = 0;
  for (int b = 0; b < nb; ++b) {
    bstart[b] = pos;
    for (int t = 0; t < threads; ++t) {
      off[t][b] = pos;
      pos += hist[t][b];
    }
  }
  bstart[nb] = pos;
  std::vector<uint64_t> tmp(n);
  parallel_chunks(n, threads, [&](int t, int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) tmp[off[t][keys[i] / width]++] = keys[i];
  });
  std::vector<int64_t> uniq(nb, 0);
  parallel_for(nb, threads, [&](int64_t lo, int64_t hi) {
    for (int64_t b = lo; b < hi; ++b) {
      auto first = tmp.begin() + bstart[b], last = tmp.begin() + bstart[b + 1];
      std::sort(first, last);
      uniq[b] = std::unique(first, last) - first;
    }
  });
  int64_t w = 0;
  for (int b = 0; b < nb; ++b) {
    std::copy(tmp.begin() + bstart[b], tmp.begin() + bstart[b] + uniq[b], keys.begin() + w);
    w += uniq[b];
  }
  keys.resize(w);
}

}  // namespace mmsb
