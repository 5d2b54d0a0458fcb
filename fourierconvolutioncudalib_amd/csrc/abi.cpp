// abi.cpp -- the exported C ABI: the reference's eleven symbols
// (src/convolution3Dfft.h:29-84) and the additive fcamd_* entry points.
// No exception or exit() crosses this boundary.
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/convolution3Dfft.h"
#include "../../include/fourierconvolution_amd.h"
#include "conv_driver.hpp"
#include "slabs.hpp"
#include "staging.hpp"
#include "static_kernels.hpp"

using namespace fcamd;

namespace fcamd {

namespace {
thread_local int t_code = 0;
thread_local std::string t_msg;
}  // namespace

void set_error(int code, const std::string& msg) {
  t_code = code;
  t_msg = msg;
}
void clear_error() {
  t_code = 0;
  t_msg.clear();
}
int last_error_code() { return t_code; }
const char* last_error_message() { return t_msg.c_str(); }

}  // namespace fcamd

namespace {

template <class F>
int guarded(F&& f) {
  clear_error();
  try {
    f();
    return kOk;
  } catch (const Error& e) {
    set_error(e.code, e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error(kOutOfMemory, "host allocation failed");
    return kOutOfMemory;
  } catch (...) {
    set_error(kInternal, "unexpected internal error");
    return kInternal;
  }
}

ConvShape shape_from(const int* imDim, const int* kernelDim) {
  FC_REQUIRE(imDim != nullptr && kernelDim != nullptr, kInvalidArgument, "null dimension array");
  ConvShape s{};
  s.nx = imDim[0];
  s.ny = imDim[1];
  s.nz = imDim[2];
  s.kx = kernelDim[0];
  s.ky = kernelDim[1];
  s.kz = kernelDim[2];
  s.cx = s.kx / 2;
  s.cy = s.ky / 2;
  s.cz = s.kz / 2;
  validate_shape(s);
  return s;
}

// shape_from for the entry points that also take the long-axis path
ConvShape shape_from_any(const int* imDim, const int* kernelDim) {
  FC_REQUIRE(imDim != nullptr && kernelDim != nullptr, kInvalidArgument, "null dimension array");
  ConvShape s{imDim[0], imDim[1], imDim[2], kernelDim[0], kernelDim[1], kernelDim[2],
              kernelDim[0] / 2, kernelDim[1] / 2, kernelDim[2] / 2};
  if (needs_long_path(s))
    validate_long_shape(s);
  else
    validate_shape(s);
  return s;
}

// device workspace of one device-resident convolution of `s`
size_t conv_workspace_bytes(const ConvShape& s) {
  return needs_long_path(s) ? long_workspace_bytes(s) : workspace_layout(s).total();
}

void enqueue_any(const ConvShape& s, float* d_out, const float* d_image, const float* d_kernel,
                 Op op, void* ws, hipStream_t st) {
  if (needs_long_path(s))
    enqueue_convolution_long(s, d_out, d_image, d_kernel, op, ws, st);
  else
    enqueue_convolution(s, d_out, d_image, d_kernel, op, ws, st);
}

// the drop-in kernel centring without the kernel <= image requirement (the
// boundary-mode paths validate the padded transform shape instead)
ConvShape raw_shape(const int* imDim, const int* kernelDim) {
  FC_REQUIRE(imDim != nullptr && kernelDim != nullptr, kInvalidArgument, "null dimension array");
  ConvShape s{imDim[0], imDim[1], imDim[2], kernelDim[0], kernelDim[1], kernelDim[2],
              kernelDim[0] / 2, kernelDim[1] / 2, kernelDim[2] / 2};
  FC_REQUIRE(s.nx >= 1 && s.ny >= 1 && s.nz >= 1 && s.kx >= 1 && s.ky >= 1 && s.kz >= 1,
             kInvalidArgument, "dimensions must be positive");
  return s;
}

Op op_from(int op) {
  FC_REQUIRE(op == FCAMD_CONVOLVE || op == FCAMD_CORRELATE, kInvalidArgument, "unknown op");
  return op == FCAMD_CORRELATE ? Op::kCorrelate : Op::kConvolve;
}

// Host-pointer convolution on device `dev`: H2D, passes, D2H into `out`.
// `out` may alias `im`.  Serialised per device (pooled buffers).
// z-slab streaming of large host-pointer calls; FCAMD_NO_SLABS=1 turns it
// off (single-shot upload, convolution, download)
bool slabs_enabled() {
  static const bool on = [] {
    const char* v = std::getenv("FCAMD_NO_SLABS");
    return !(v && v[0] && v[0] != '0');
  }();
  return on;
}

// Device memory the host-pointer path may use on `dev`: FCAMD_DEVICE_BUDGET_MB
// when set (read on every call), else the device's free memory plus what the
// pool already holds, less a 256 MiB margin.
bool env_budget(size_t* bytes) {
  const char* v = std::getenv("FCAMD_DEVICE_BUDGET_MB");
  if (!v || !v[0]) return false;
  const long long mb = std::atoll(v);
  FC_REQUIRE(mb > 0, kInvalidArgument, "FCAMD_DEVICE_BUDGET_MB must be a positive integer");
  *bytes = (size_t)mb << 20;
  return true;
}

size_t device_budget(const DevicePool& pool) {
  size_t b = 0;
  if (env_budget(&b)) return b;
  size_t free_b = 0, total_b = 0;
  FC_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
  const size_t held = pool.img_bytes + pool.out_bytes + pool.ws_bytes + pool.ker_bytes;
  const size_t avail = free_b + held, margin = size_t(256) << 20;
  return avail > margin ? avail - margin : 0;
}

// How a host-pointer convolution of `s` runs within `budget`: a z-slab plan
// (nslabs > 0) or the single shot.
struct HostPlan {
  SlabPlan sp;
  size_t device_bytes = 0;  // everything it allocates on the device
};

HostPlan host_plan(const ConvShape& s, size_t budget) {
  HostPlan h;
  if (needs_long_path(s)) {  // long axes: one shot, no slabs
    h.device_bytes = (sizeof(float) * (size_t)s.voxels() + 255) / 256 * 256 +
                     long_workspace_bytes(s) +
                     (sizeof(float) * (size_t)s.kx * s.ky * s.kz + 255) / 256 * 256;
    return h;
  }
  const size_t single = single_shot_device_bytes(s);
  const size_t nker = (sizeof(float) * (size_t)s.kx * s.ky * s.kz + 255) / 256 * 256;
  if (slabs_enabled()) {
    h.sp = slab_plan(s, budget > nker ? budget - nker : 0, single > budget);
    if (h.sp.nslabs > 0) {
      h.device_bytes = h.sp.device_bytes() + nker;
      return h;
    }
  }
  h.device_bytes = single;
  return h;
}

void host_convolution(const ConvShape& s, const float* im, const float* kernel, float* out,
                      Op op, int dev) {
  FC_REQUIRE(im != nullptr && kernel != nullptr && out != nullptr, kInvalidArgument,
             "null data pointer");
  DeviceGuard guard(dev);
  DevicePool& pool = device_pool(dev);
  std::lock_guard<std::mutex> lk(pool.mu);
  if (!pool.stream) FC_HIP_CHECK(hipStreamCreateWithFlags(&pool.stream, hipStreamNonBlocking));
  const size_t nimg = sizeof(float) * (size_t)s.voxels();
  const size_t nker = sizeof(float) * (size_t)s.kx * s.ky * s.kz;
  const size_t budget = device_budget(pool);
  const HostPlan hp = host_plan(s, budget);
  if (hp.sp.nslabs > 0) {
    // z-slab streaming through a ring of slab buffers, both PCIe directions at once
    const SlabPlan& sp = hp.sp;
    void* ring = pool_reserve(reinterpret_cast<void**>(&pool.img), &pool.img_bytes,
                              sp.ring_bytes);
    void* wrap = pool_reserve(reinterpret_cast<void**>(&pool.out), &pool.out_bytes,
                              sp.wrap_bytes);
    void* dker = pool_reserve(reinterpret_cast<void**>(&pool.ker), &pool.ker_bytes, nker);
    void* ws = pool_reserve(&pool.ws, &pool.ws_bytes, sp.ws_bytes);
    host_convolution_slabs(s, sp, im, kernel, out, op, dev, static_cast<float*>(ring),
                           static_cast<float*>(wrap), static_cast<float*>(dker), ws, pool.stream,
                           &host_copy_parallel);
    return;
  }
  FC_REQUIRE(hp.device_bytes <= budget, kOutOfMemory,
             "the convolution needs " + std::to_string(hp.device_bytes >> 20) +
                 " MiB of device memory (device budget " + std::to_string(budget >> 20) +
                 " MiB) and no z-slab plan fits the budget");
  const size_t nws = conv_workspace_bytes(s);
  void* dimg = pool_reserve(reinterpret_cast<void**>(&pool.img), &pool.img_bytes, nimg);
  void* dker = pool_reserve(reinterpret_cast<void**>(&pool.ker), &pool.ker_bytes, nker);
  void* ws = pool_reserve(&pool.ws, &pool.ws_bytes, nws);
  staged_upload(dev, dimg, im, nimg, pool.stream);
  staged_upload(dev, dker, kernel, nker, pool.stream);
  enqueue_any(s, static_cast<float*>(dimg), static_cast<float*>(dimg), static_cast<float*>(dker),
              op, ws, pool.stream);
  staged_download(dev, out, dimg, nimg, pool.stream);
}

// device bytes a host-pointer convolution of `s` would allocate on `dev` now
long long host_device_bytes(const ConvShape& s, int dev) {
  size_t b = 0;
  if (env_budget(&b)) return (long long)host_plan(s, b).device_bytes;  // no device needed
  DeviceGuard guard(dev);
  DevicePool& pool = device_pool(dev);
  std::lock_guard<std::mutex> lk(pool.mu);
  return (long long)host_plan(s, device_budget(pool)).device_bytes;
}

// host-pointer boundary-mode convolution through the device pool
void host_convolution_padded(const PaddedProblem& p, const float* im, const float* kernel,
                             float* out, Op op, int dev) {
  FC_REQUIRE(im != nullptr && kernel != nullptr && out != nullptr, kInvalidArgument,
             "null data pointer");
  DeviceGuard guard(dev);
  DevicePool& pool = device_pool(dev);
  std::lock_guard<std::mutex> lk(pool.mu);
  if (!pool.stream) FC_HIP_CHECK(hipStreamCreateWithFlags(&pool.stream, hipStreamNonBlocking));
  const ConvShape& s = p.padded;
  const size_t nimg = sizeof(float) * (size_t)p.vp.n[0] * p.vp.n[1] * p.vp.n[2];
  const size_t nker = sizeof(float) * (size_t)s.kx * s.ky * s.kz;
  const size_t nws = workspace_layout(s).total();
  void* dimg = pool_reserve(reinterpret_cast<void**>(&pool.img), &pool.img_bytes, nimg);
  void* dker = pool_reserve(reinterpret_cast<void**>(&pool.ker), &pool.ker_bytes, nker);
  void* ws = pool_reserve(&pool.ws, &pool.ws_bytes, nws);
  staged_upload(dev, dimg, im, nimg, pool.stream);
  staged_upload(dev, dker, kernel, nker, pool.stream);
  enqueue_convolution_padded(p, static_cast<float*>(dimg), static_cast<float*>(dimg),
                             static_cast<float*>(dker), op, ws, pool.stream);
  staged_download(dev, out, dimg, nimg, pool.stream);
}

}  // namespace

// ===========================================================================
// reference symbols
// ===========================================================================
extern "C" {

// src/convolution3Dfft.h:56, src/convolution3Dfft.cu:399-578
void convolution3DfftCUDAInPlace(imageType* im, int* imDim, imageType* kernel, int* kernelDim,
                                 int devCUDA) {
  (void)guarded([&] {
    ConvShape s = shape_from_any(imDim, kernelDim);
    host_convolution(s, im, kernel, im, Op::kConvolve, devCUDA);
  });
}

// src/convolution3Dfft.h:41-45, src/convolution3Dfft.cu:299-391.  Dims are
// z-fastest (imDim[2] contiguous), so {imDim[2], imDim[1], imDim[0]} is the
// {x, y, z} of the in-place entry point.  The reference computes this
// in-place on an unpadded R2C buffer, valid only under the long-removed
// cuFFT native layout (:298,358-361); here the call returns the convolution.
imageType* convolution3DfftCUDA(imageType* im, int* imDim, imageType* kernel, int* kernelDim,
                                int devCUDA) {
  float* result = nullptr;
  int rc = guarded([&] {
    FC_REQUIRE(imDim != nullptr && kernelDim != nullptr, kInvalidArgument,
               "null dimension array");
    const int d[3] = {imDim[2], imDim[1], imDim[0]};
    const int k[3] = {kernelDim[2], kernelDim[1], kernelDim[0]};
    ConvShape s = shape_from_any(d, k);
    result = new float[(size_t)s.voxels()];
    host_convolution(s, im, kernel, result, Op::kConvolve, devCUDA);
  });
  if (rc != kOk) {
    delete[] result;
    return nullptr;
  }
  return result;
}

// src/convolution3Dfft.h:29-32, src/convolution3Dfft.cu:216-292: the kernel is
// image-sized and used without the fftShift (its tap 0 sits at index 0).
imageType* convolution3DfftCUDA_test(imageType* im, int* imDim, imageType* kernel, int devCUDA) {
  float* result = nullptr;
  int rc = guarded([&] {
    FC_REQUIRE(imDim != nullptr, kInvalidArgument, "null dimension array");
    ConvShape s{};
    s.nx = imDim[2];
    s.ny = imDim[1];
    s.nz = imDim[0];
    s.kx = s.nx;
    s.ky = s.ny;
    s.kz = s.nz;
    s.cx = s.cy = s.cz = 0;
    if (needs_long_path(s))
      validate_long_shape(s);
    else
      validate_shape(s);
    result = new float[(size_t)s.voxels()];
    host_convolution(s, im, kernel, result, Op::kConvolve, devCUDA);
  });
  if (rc != kOk) {
    delete[] result;
    return nullptr;
  }
  return result;
}

// src/convolution3Dfft.cu:581-584
int cuda_version() { return HIP_VERSION; }

// src/convolution3Dfft.cu:587-603 (cufftEstimate*d, the FFT work area of the
// reference's path).  Here: the device memory a host-pointer convolution of
// an image of `shape` (len 1..3, last entry contiguous) allocates on the
// current device under its device budget (z-slab ring or single shot), for a
// one-voxel kernel, in MiB rounded up; fcamd_host_device_bytes gives the
// figure for a given kernel.  -1 for an invalid request.
int gpu_mem_needed_mb(int* shape, int len) {
  if (shape == nullptr || len < 1 || len > 3) return -1;
  int d[3] = {1, 1, 1};  // {x, y, z}
  for (int i = 0; i < len; ++i) {
    if (shape[i] < 1) return -1;
    d[len - 1 - i] = shape[i];
  }
  const int k[3] = {1, 1, 1};
  long long b = -1;
  const int rc = guarded([&] {
    int dev = 0;
    size_t env = 0;
    if (!env_budget(&env)) FC_HIP_CHECK(hipGetDevice(&dev));
    b = host_device_bytes(shape_from_any(d, k), dev);
  });
  if (rc != kOk || b < 0) return -1;
  return (int)((b + (1 << 20) - 1) >> 20);
}

// ===========================================================================
// src/standardCUDAfunctions.cu:13-71 on HIP
// ===========================================================================
int getNumDevicesCUDA() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int selectDeviceWithHighestComputeCapability() {
  const int n = getNumDevicesCUDA();
  int best = -1, best_cc = 0;
  for (int d = 0; d < n; ++d) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, d) != hipSuccess) continue;
    const int cc = 10 * p.major + p.minor;
    if (cc > best_cc) {
      best_cc = cc;
      best = d;
    }
  }
  return best;
}

int getCUDAcomputeCapabilityMajorVersion(int devCUDA) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, devCUDA) != hipSuccess) return -1;
  return p.major;
}

int getCUDAcomputeCapabilityMinorVersion(int devCUDA) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, devCUDA) != hipSuccess) return -1;
  return p.minor;
}

void getNameDeviceCUDA(int devCUDA, char* name) {
  if (name == nullptr) return;
  char buf[256];
  std::memset(buf, 0, sizeof(buf));
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, devCUDA) == hipSuccess) {
    // the marketing name can be empty (no name table on the host); the
    // architecture name (e.g. gfx950:sramecc+:xnack-) is always set
    std::strncpy(buf, p.name[0] ? p.name : p.gcnArchName, sizeof(buf) - 1);
  }
  std::memcpy(name, buf, sizeof(buf));
}

long long int getMemDeviceCUDA(int devCUDA) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, devCUDA) != hipSuccess) return -1;
  return (long long)p.totalGlobalMem;
}

// ===========================================================================
// additive entry points
// ===========================================================================
int fcamd_last_error(void) { return last_error_code(); }
const char* fcamd_last_error_message(void) { return last_error_message(); }
void fcamd_free_result(float* p) { delete[] p; }

int fcamd_convolve3d_host(float* im, const int* imDim, const float* kernel, const int* kernelDim,
                          int op, int devCUDA) {
  return guarded([&] {
    ConvShape s = shape_from_any(imDim, kernelDim);
    host_convolution(s, im, kernel, im, op_from(op), devCUDA);
  });
}

long long fcamd_host_device_bytes(const int* imDim, const int* kernelDim, int devCUDA) {
  long long r = -1;
  (void)guarded([&] { r = host_device_bytes(shape_from_any(imDim, kernelDim), devCUDA); });
  return r;
}

long long fcamd_workspace_bytes(const int* imDim, const int* kernelDim) {
  long long r = -1;
  (void)guarded([&] { r = (long long)conv_workspace_bytes(shape_from_any(imDim, kernelDim)); });
  return r;
}

int fcamd_convolve3d_device(float* d_out, const float* d_image, const int* imDim,
                            const float* d_kernel, const int* kernelDim, int op, void* d_workspace,
                            long long workspace_bytes, void* stream) {
  return guarded([&] {
    FC_REQUIRE(d_out && d_image && d_kernel, kInvalidArgument, "null device pointer");
    ConvShape s = shape_from_any(imDim, kernelDim);
    const Op o = op_from(op);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t need = conv_workspace_bytes(s);
    if (d_workspace) {
      FC_REQUIRE(workspace_bytes >= (long long)need, kInvalidArgument, "workspace too small");
      enqueue_any(s, d_out, d_image, d_kernel, o, d_workspace, st);
    } else {
      int dev = 0;
      FC_HIP_CHECK(hipGetDevice(&dev));
      DevicePool& pool = device_pool(dev);
      std::lock_guard<std::mutex> lk(pool.mu);
      void* ws = pool_reserve(&pool.ws, &pool.ws_bytes, need);
      enqueue_any(s, d_out, d_image, d_kernel, o, ws, st);
      // the pooled workspace is reused by the next call: finish before unlocking
      FC_HIP_CHECK(hipStreamSynchronize(st));
    }
  });
}

long long fcamd_boundary_workspace_bytes(const int* imDim, const int* kernelDim, int boundary) {
  long long r = -1;
  (void)guarded([&] {
    r = (long long)workspace_layout(padded_problem(raw_shape(imDim, kernelDim), boundary).padded)
            .total();
  });
  return r;
}

int fcamd_padded_extents(const int* imDim, const int* kernelDim, int boundary, int* padded) {
  return guarded([&] {
    FC_REQUIRE(padded != nullptr, kInvalidArgument, "null output array");
    const PaddedProblem p = padded_problem(raw_shape(imDim, kernelDim), boundary);
    padded[0] = p.padded.nx;
    padded[1] = p.padded.ny;
    padded[2] = p.padded.nz;
  });
}

int fcamd_convolve3d_boundary_device(float* d_out, const float* d_image, const int* imDim,
                                     const float* d_kernel, const int* kernelDim, int op,
                                     int boundary, void* d_workspace, long long workspace_bytes,
                                     void* stream) {
  return guarded([&] {
    FC_REQUIRE(d_out && d_image && d_kernel, kInvalidArgument, "null device pointer");
    const PaddedProblem p = padded_problem(raw_shape(imDim, kernelDim), boundary);
    const Op o = op_from(op);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t need = workspace_layout(p.padded).total();
    if (d_workspace) {
      FC_REQUIRE(workspace_bytes >= (long long)need, kInvalidArgument, "workspace too small");
      enqueue_convolution_padded(p, d_out, d_image, d_kernel, o, d_workspace, st);
    } else {
      int dev = 0;
      FC_HIP_CHECK(hipGetDevice(&dev));
      DevicePool& pool = device_pool(dev);
      std::lock_guard<std::mutex> lk(pool.mu);
      void* ws = pool_reserve(&pool.ws, &pool.ws_bytes, need);
      enqueue_convolution_padded(p, d_out, d_image, d_kernel, o, ws, st);
      // the pooled workspace is reused by the next call: finish before unlocking
      FC_HIP_CHECK(hipStreamSynchronize(st));
    }
  });
}

int fcamd_convolve3d_boundary_host(float* im, const int* imDim, const float* kernel,
                                   const int* kernelDim, int op, int boundary, int devCUDA) {
  return guarded([&] {
    const PaddedProblem p = padded_problem(raw_shape(imDim, kernelDim), boundary);
    host_convolution_padded(p, im, kernel, im, op_from(op), devCUDA);
  });
}

long long fcamd_grad_kernel_workspace_bytes(const int* imDim, const int* kernelDim) {
  long long r = -1;
  (void)guarded([&] {
    const ConvShape s = shape_from_any(imDim, kernelDim);
    r = (long long)(needs_long_path(s) ? long_workspace_bytes(s) : grad_workspace_layout(s).total());
  });
  return r;
}

int fcamd_conv_grad_kernel_device(float* d_dkernel, const float* d_grad_out, const float* d_image,
                                  const int* imDim, const int* kernelDim, void* d_workspace,
                                  long long workspace_bytes, void* stream) {
  return guarded([&] {
    FC_REQUIRE(d_dkernel && d_grad_out && d_image, kInvalidArgument, "null device pointer");
    ConvShape s = shape_from_any(imDim, kernelDim);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const bool lng = needs_long_path(s);
    const size_t need = lng ? long_workspace_bytes(s) : grad_workspace_layout(s).total();
    auto run = [&](void* ws) {
      if (lng)
        enqueue_grad_kernel_long(s, d_dkernel, d_grad_out, d_image, ws, st);
      else
        enqueue_grad_kernel(s, d_dkernel, d_grad_out, d_image, ws, st);
    };
    if (d_workspace) {
      FC_REQUIRE(workspace_bytes >= (long long)need, kInvalidArgument, "workspace too small");
      run(d_workspace);
    } else {
      int dev = 0;
      FC_HIP_CHECK(hipGetDevice(&dev));
      DevicePool& pool = device_pool(dev);
      std::lock_guard<std::mutex> lk(pool.mu);
      void* ws = pool_reserve(&pool.ws, &pool.ws_bytes, need);
      run(ws);
      FC_HIP_CHECK(hipStreamSynchronize(st));
    }
  });
}

void fcamd_pass_timing(int enable) { pass_timing_enable(enable != 0); }

int fcamd_pass_times(double* ms, int* counts, int cap) {
  double m[kNumPasses];
  int c[kNumPasses];
  int rc = guarded([&] { pass_timing_collect(m, c); });
  if (rc != kOk) return -rc;
  for (int i = 0; i < kNumPasses && i < cap; ++i) {
    ms[i] = m[i];
    counts[i] = c[i];
  }
  return kNumPasses;
}

const char* fcamd_pass_name(int pass) { return pass_name(pass); }

int fcamd_radix_plan(int length, int* radices, int cap) {
  if (length < 1) return -1;
  std::vector<int> r = radix_plan(length);
  for (int i = 0; i < (int)r.size() && i < cap; ++i) radices[i] = r[i];
  return (int)r.size();
}

}  // extern "C"
