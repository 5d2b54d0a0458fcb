// noalloc.hip -- fcamd_convolve3d_device_noalloc: the device-resident
// convolution with every scratch buffer in one file-scope device array, for
// callers that cannot allocate (code under graph capture, tight serving
// loops, harnesses that forbid allocation).
//
// The array `scratch` holds the workspace of fcamd_convolve3d_device (the
// half spectrum and the kernel's rows and planes, at most
// FCAMD_NOALLOC_MAX_WORKSPACE bytes) followed by the twiddle tables of the
// call's three transform lengths.  The tables are generated on the device
// by a small kernel from values passed by launch argument, so a call
// allocates nothing, copies nothing from host memory and can be recorded
// into a graph.  Each device holds its own copy of the array (taken when
// the library's code is loaded on it) until the process ends; its address
// is looked up once per device with hipGetSymbolAddress and only that
// address is handed to kernels.  Every call on a device shares the array:
// calls that overlap on one device race.
#include <mutex>

#include "../../include/fourierconvolution_amd.h"
#include "conv_driver.hpp"

namespace fcamd {
namespace {

constexpr size_t kWorkspaceMax = (size_t)FCAMD_NOALLOC_MAX_WORKSPACE;
constexpr size_t kTableBytes = size_t(1) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kWorkspaceMax + kTableBytes];

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// scratch's address on the current device (looked up on the first call)
char* scratch_address(int dev) {
  FC_REQUIRE(dev >= 0 && dev < kMaxDevices, kInvalidArgument, "device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    FC_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    FC_REQUIRE(p != nullptr, kHipError, "no device address for the scratch array");
    g_scratch[dev] = p;
  }
  return static_cast<char*>(g_scratch[dev]);
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Twiddle tables of one length: tw[j] = exp(-2 pi i j / L), j < L, and the
// compile-time plan's stage rows (layout of stage_table_segments), each
// angle in FP64 and rounded once to FP32 like the host-built tables.
struct TableJob {
  float2* tw;
  float2* stw;
  int L;
  int nseg;
  int stw_total;
  TwSegment seg[kMaxTwSegments];
};

__global__ void __launch_bounds__(256) generate_tables(TableJob job) {
  const int n = job.L + job.stw_total;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    long m, span;
    float2* dst;
    if (i < job.L) {
      m = i;
      span = job.L;
      dst = job.tw + i;
    } else {
      const int e = i - job.L;
      int s = 0;
      while (s + 1 < job.nseg && job.seg[s + 1].offset <= e) ++s;
      const TwSegment g = job.seg[s];
      const int k = (e - g.offset) / g.rp, r = (e - g.offset) % g.rp;
      dst = job.stw + e;
      if (r >= g.r) {
        *dst = make_float2(1.f, 0.f);
        continue;
      }
      span = (long)g.ns * g.r;
      m = ((long)r * k) % span;
    }
    // exp(-2 pi i m / span) = cos(pi x) - i sin(pi x), x = 2 m / span
    double sn, cs;
    sincospi(2.0 * (double)m / (double)span, &sn, &cs);
    *dst = make_float2((float)cs, (float)-sn);
  }
}

// prepares the FftDesc of length L with its tables at `base`; returns the
// bytes used
size_t prepare_length(int L, char* base, size_t room, FftDesc* d, hipStream_t st) {
  TableJob job{};
  job.L = L;
  job.nseg = stage_table_segments(L, job.seg, &job.stw_total);
  FC_REQUIRE(job.nseg <= kMaxTwSegments, kInternal, "too many twiddle stages");
  const size_t tw_bytes = align256(sizeof(float2) * (size_t)L);
  const size_t stw_bytes = align256(sizeof(float2) * (size_t)job.stw_total);
  FC_REQUIRE(tw_bytes + stw_bytes <= room, kUnsupportedSize, "twiddle tables exceed the scratch");
  job.tw = reinterpret_cast<float2*>(base);
  job.stw = job.stw_total > 0 ? reinterpret_cast<float2*>(base + tw_bytes) : nullptr;
  const int n = L + job.stw_total;
  const unsigned blocks = (unsigned)((n + 255) / 256 < 64 ? (n + 255) / 256 : 64);
  hipLaunchKernelGGL(generate_tables, dim3(blocks), dim3(256), 0, st, job);
  *d = FftDesc{};
  d->L = L;
  d->nst = radix_plan_into(L, d->rad, kMaxStages);
  FC_REQUIRE(d->nst <= kMaxStages, kUnsupportedSize, "FFT plan too deep");
  d->tw = job.tw;
  d->stw = job.stw;
  return tw_bytes + stw_bytes;
}

// clears the thread's table override however the call ends
struct OverrideScope {
  explicit OverrideScope(const FftDescSet* s) { set_fft_desc_override(s); }
  ~OverrideScope() { set_fft_desc_override(nullptr); }
};

template <class F>
int noalloc_guarded(F&& f) {
  clear_error();
  try {
    f();
    return kOk;
  } catch (const Error& e) {
    set_error(e.code, e.msg);
    return e.code;
  } catch (...) {
    set_error(kInternal, "unexpected internal error");
    return kInternal;
  }
}

}  // namespace
}  // namespace fcamd

using namespace fcamd;

extern "C" int fcamd_convolve3d_device_noalloc(float* d_out, const float* d_image,
                                               const int* imDim, const float* d_kernel,
                                               const int* kernelDim, int op, void* stream) {
  return noalloc_guarded([&] {
    FC_REQUIRE(d_out && d_image && d_kernel, kInvalidArgument, "null device pointer");
    FC_REQUIRE(imDim && kernelDim, kInvalidArgument, "null dimension array");
    FC_REQUIRE(op == FCAMD_CONVOLVE || op == FCAMD_CORRELATE, kInvalidArgument, "unknown op");
    ConvShape s{imDim[0], imDim[1], imDim[2], kernelDim[0], kernelDim[1], kernelDim[2],
                kernelDim[0] / 2, kernelDim[1] / 2, kernelDim[2] / 2};
    validate_shape(s);
    const size_t need = workspace_layout(s).total();
    FC_REQUIRE(need <= kWorkspaceMax, kUnsupportedSize,
               "the convolution needs " + std::to_string(need) +
                   " bytes of workspace; the allocation-free entry point holds at most " +
                   std::to_string(kWorkspaceMax));
    hipStream_t st = static_cast<hipStream_t>(stream);
    int dev = 0;
    FC_HIP_CHECK(hipGetDevice(&dev));
    char* base = scratch_address(dev);
    driver_init();
    FftDescSet set;
    char* tables = base + kWorkspaceMax;
    size_t used = 0;
    int lens[5];
    const int nlen = table_lengths(s, lens);
    for (int i = 0; i < nlen; ++i) {
      used += prepare_length(lens[i], tables + used, kTableBytes - used, &set.d[set.n], st);
      ++set.n;
    }
    OverrideScope scope(&set);
    enqueue_convolution(s, d_out, d_image, d_kernel,
                        op == FCAMD_CORRELATE ? Op::kCorrelate : Op::kConvolve, base, st);
  });
}
