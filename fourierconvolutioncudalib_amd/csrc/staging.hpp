// staging.hpp -- host<->device copies of pageable host memory through pinned
// staging buffers.
//
// A pageable copy runs at about half the link rate (the runtime stages it
// itself, one chunk at a time).  Here the transfer is cut into chunks that
// cycle through a few pinned buffers: the host-side copy of chunk i+1
// (split over several threads) overlaps the DMA of chunk i, and the DMA
// engines only ever see pinned memory.
#pragma once

#include <cstddef>

#include "fc_internal.hpp"

namespace fcamd {

// Enqueues the upload of `bytes` from pageable `h_src` to `d_dst` on `st`
// and returns once every chunk has been handed to the DMA engine (the
// source may then be reused; the copy completes in stream order).
void staged_upload(int dev, void* d_dst, const void* h_src, size_t bytes, hipStream_t st);
// Copies `bytes` from `d_src` (after the work already on `st`) to pageable
// `h_dst`; returns when the data is in `h_dst`.
void staged_download(int dev, void* h_dst, const void* d_src, size_t bytes, hipStream_t st);
// host memcpy split over the staging copy threads
void host_copy_parallel(void* dst, const void* src, size_t bytes);

}  // namespace fcamd
