// slabs.hpp -- z-slab (overlap-save) streaming of host-pointer convolutions
// (slabs.cpp).
#pragma once

#include "conv_driver.hpp"

namespace fcamd {

// volumes below this size take the single-shot host path
constexpr size_t kSlabMinBytes = size_t(32) << 20;

struct SlabPlan {
  int P = 0;       // padded z length of every slab
  int m = 0;       // output planes per slab (the last may have fewer)
  int nslabs = 0;  // 0: slabs do not apply
};

SlabPlan slab_plan(const ConvShape& s);
// the padded problem of output planes [z0, z0 + m), transformed at z length P
PaddedProblem slab_problem(const ConvShape& s, int z0, int m, int P);
size_t slab_workspace_bytes(const ConvShape& s, const SlabPlan& sp);

// Convolves host `im` into host `out` (may alias) slab by slab.  d_in and
// d_out hold one volume each, d_ker the kernel, `ws` slab_workspace_bytes;
// `compute` runs the passes, `host_copy` does the pageable<->pinned copies.
void host_convolution_slabs(const ConvShape& s, const SlabPlan& sp, const float* im,
                            const float* kernel, float* out, Op op, int dev, float* d_in,
                            float* d_out, float* d_ker, void* ws, hipStream_t compute,
                            void (*host_copy)(void*, const void*, size_t));

}  // namespace fcamd
