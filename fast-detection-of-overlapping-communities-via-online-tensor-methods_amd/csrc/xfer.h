// Host <-> device copies of large pageable buffers through a pinned, double-buffered staging area
// (one per process, freed by release_staging): the host copies one chunk into pinned memory with several threads while the
// DMA engine moves the previous chunk. Small copies go straight through hipMemcpyAsync.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>

namespace mmsb {

// Copies `bytes` from host `src` to device `dst` on stream `s`; returns once `src` may be reused
// (the last chunk may still be in flight on `s`).
int h2d(hipStream_t s, void* dst, const void* src, size_t bytes);
// Copies `bytes` from device `src` to host `dst` on stream `s` (after the work queued on `s`);
// returns when `dst` holds the data.
int d2h(hipStream_t s, void* dst, const void* src, size_t bytes);
// Frees the pinned staging buffers and their events (re-created by the next large copy).
void release_staging();
// Frees the pinned staging, device slots, events and copy stream of the host-CSR SpMM
// (spmm_stream.cpp), re-created by its next call.
void release_streamer();
// Frees the device temporaries kept by the CSR transposes and row orderings (csr_transpose.hip).
void release_sort_scratch();

}  // namespace mmsb
