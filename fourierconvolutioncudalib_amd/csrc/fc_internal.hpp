// fc_internal.hpp -- shared host/device definitions of the MI355X 3D FFT
// convolution library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

namespace fcamd {

// ---------------------------------------------------------------------------
// error state: every C entry point returns a code and records a message in a
// thread-local slot; nothing throws or exits across the C ABI.
// ---------------------------------------------------------------------------
enum Status : int {
  kOk = 0,
  kInvalidArgument = 1,
  kUnsupportedSize = 2,
  kOutOfMemory = 3,
  kHipError = 4,
  kNoDevice = 5,
  kInternal = 6,
};

void set_error(int code, const std::string& msg);
void clear_error();
int last_error_code();
const char* last_error_message();

struct Error {
  int code;
  std::string msg;
};

#define FC_HIP_CHECK(expr)                                                    \
  do {                                                                        \
    hipError_t fc_e_ = (expr);                                                \
    if (fc_e_ != hipSuccess)                                                  \
      throw ::fcamd::Error{fc_e_ == hipErrorOutOfMemory ? ::fcamd::kOutOfMemory \
                                                        : ::fcamd::kHipError,   \
                           std::string(#expr) + ": " + hipGetErrorString(fc_e_)}; \
  } while (0)

#define FC_REQUIRE(cond, code, msg)                                           \
  do {                                                                        \
    if (!(cond)) throw ::fcamd::Error{(code), (msg)};                         \
  } while (0)

// ---------------------------------------------------------------------------
// sizes
// ---------------------------------------------------------------------------
// Complex rows of the half spectrum are padded to a multiple of this many
// bins so every 16-bin column tile starts on a 128-byte line.
constexpr int kRowAlign = 16;

inline int half_bins(int nx) { return nx / 2 + 1; }
inline int padded_bins(int nx) {
  return (half_bins(nx) + kRowAlign - 1) / kRowAlign * kRowAlign;
}

// Largest radix stage count a plan may hold.
constexpr int kMaxStages = 24;

// Device view of one 1D transform length: radix sequence (Stockham order) and
// the FP64-generated, FP32-rounded table tw[j] = exp(-2*pi*i*j/L), j < L.
// With M > 0 the length-L transform runs as Bluestein's chirp-z: a circular
// convolution of length M >= 2L - 1 (7-smooth); rad/nst/tw are then M's
// plan and table, chirp[n] = exp(-i pi n^2 / L) for n < L, and bfft the
// length-M spectrum of the conjugate chirp sequence, scaled by 1/M.
struct FftDesc {
  int L;
  int nst;
  int rad[kMaxStages];
  const float2* tw;   // runtime-length kernels
  const float2* stw;  // per-stage tables of the compile-time plan (or null)
  int M;              // 0: plain Stockham plan of L
  const float2* chirp;
  const float2* bfft;
  // positions a transform occupies in a tile buffer
  __host__ __device__ int buf() const { return M > 0 ? M : L; }
};

}  // namespace fcamd
