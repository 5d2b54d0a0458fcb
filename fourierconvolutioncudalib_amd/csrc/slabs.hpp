// slabs.hpp -- z-slab (overlap-save) streaming of host-pointer convolutions
// through a ring of slab buffers (slabs.cpp).
#pragma once

#include "conv_driver.hpp"

namespace fcamd {

// volumes below this size take the single-shot host path unless it does not
// fit the device budget
constexpr size_t kSlabMinBytes = size_t(32) << 20;
// slab input buffers in the ring: slab k uploads into one while slab k-1
// computes in another and slab k-2 is downloaded from the third
constexpr int kRingBuffers = 3;

struct SlabPlan {
  int P = 0;       // padded z length of every slab's transform
  int m = 0;       // output planes per slab (the last may have fewer)
  int nslabs = 0;  // 0: slabs do not apply
  size_t ring_bytes = 0;    // kRingBuffers windows of m + kz - 1 planes
  size_t wrap_bytes = 0;    // the first kz/2 input planes, kept for the last slab
  size_t ws_bytes = 0;      // transform workspace of one slab
  size_t device_bytes() const { return ring_bytes + wrap_bytes + ws_bytes; }
};

// Device bytes of the single-shot host path: image, workspace and kernel.
size_t single_shot_device_bytes(const ConvShape& s);

// Slab plan for a host-pointer convolution whose device buffers (ring,
// wrap planes, workspace; the kernel comes on top) must fit `budget` bytes.
// `force` plans slabs even for volumes below kSlabMinBytes.  nslabs = 0 when
// no slab length fits or the volume is too short for four slabs.
SlabPlan slab_plan(const ConvShape& s, size_t budget, bool force = false);

// The padded problem of output planes [z0, z0 + m) reading the full
// volume's planes (z0 - lo + p) mod nz, transformed at z length P.
PaddedProblem slab_problem(const ConvShape& s, int z0, int m, int P);
// The same slab reading its window from a buffer that holds exactly the
// window's m + kz - 1 planes (plane p of the buffer = window position p).
PaddedProblem slab_window_problem(const ConvShape& s, int m, int P);

// Convolves host `im` into host `out` (may alias) slab by slab.  `ring`
// holds sp.ring_bytes, `wrap` sp.wrap_bytes, `ws` sp.ws_bytes, `d_ker` the
// kernel; `compute` runs the passes, `host_copy` does the pageable<->pinned
// copies.  No full volume is ever held on the device.
void host_convolution_slabs(const ConvShape& s, const SlabPlan& sp, const float* im,
                            const float* kernel, float* out, Op op, int dev, float* ring,
                            float* wrap, float* d_ker, void* ws, hipStream_t compute,
                            void (*host_copy)(void*, const void*, size_t));

}  // namespace fcamd
