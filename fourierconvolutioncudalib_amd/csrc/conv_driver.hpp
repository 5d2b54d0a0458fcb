// conv_driver.hpp -- host driver of the 3D FFT convolution (plans, tables,
// workspaces, pass sequence).  All sizes use the reference's {x, y, z}
// naming with x the contiguous axis (src/convolution3Dfft.cu:394-421).
#pragma once

#include <mutex>
#include <vector>

#include "fc_internal.hpp"
#include "kernels.hpp"

namespace fcamd {

// One convolution problem.  kc* are the kernel centres (where kernel tap
// (0,0,0) is taken from): k/2 for the drop-in entry points, 0 for the
// pre-shifted image-sized kernel of convolution3DfftCUDA_test.
struct ConvShape {
  int nx, ny, nz;
  int kx, ky, kz;
  int cx, cy, cz;
  long long voxels() const { return (long long)nx * ny * nz; }
};

// Workspace carve-up (bytes, each region 256-byte aligned).
struct WorkspaceLayout {
  size_t spectrum;  // S   [nz][ny][nxp] float2
  size_t krows;     // Kx  [kz][ky][nxp] float2
  size_t kplanes;   // Kxy [kz][ny][nxp] float2
  size_t total() const { return spectrum + krows + kplanes; }
};

WorkspaceLayout workspace_layout(const ConvShape& s);

// Validates the shape; throws Error on failure.
void validate_shape(const ConvShape& s);

// One-time kernel attribute set-up (safe to call repeatedly).
void driver_init();

// FFT description for length L on the current device.  Its twiddle tables
// are created on first use and cached for the process; with `st` given, a
// first use while `st` is capturing a graph is refused (kInvalidArgument).
FftDesc fft_desc(int L, hipStream_t st = nullptr);

// Radix plan used for length L.
std::vector<int> radix_plan(int L);
// The same plan written into rad[0..cap) without allocating; returns the
// stage count (which may exceed cap).
int radix_plan_into(int L, int* rad, int cap);

// Layout of the per-stage twiddle table of the compile-time plan of L
// (forward stage order, then reversed): one segment per stage, `ns` rows of
// `rp` entries from element `offset` on, row k entry r = exp(-2 pi i
// (r k mod ns R) / (ns R)) for r < R and 1 for R <= r < rp.  Returns the
// segment count (0: L has no compile-time plan) and the total entries.
struct TwSegment {
  int offset, ns, r, rp;
};
constexpr int kMaxTwSegments = 32;
int stage_table_segments(int L, TwSegment* seg, int* total);

// While set on the calling thread, fft_desc() serves the lengths held here
// (tables in caller-provided device memory) and creates no table; any other
// length is an internal error.  nullptr clears it.
struct FftDescSet {
  int n = 0;
  FftDesc d[5];  // table_lengths() of one shape
};
void set_fft_desc_override(const FftDescSet* set);

enum class Op : int { kConvolve = 0, kCorrelate = 1 };

// Passes of one convolution, in launch order (for per-pass event timing).
enum Pass : int {
  kPassKernelRows = 0,
  kPassKernelPlanes,
  kPassXForward,
  kPassYForward,
  kPassZFused,
  kPassYInverse,
  kPassXInverse,
  kNumPasses
};
const char* pass_name(int p);

// Optional HIP-event timing of every pass launch (off by default).
void pass_timing_enable(bool on);
bool pass_timing_enabled();
// Waits for the recorded events and returns the summed milliseconds and the
// launch count per pass since the last call (then resets).
void pass_timing_collect(double* ms, int* count);

// Enqueues the full convolution on `stream`: d_out may alias d_image.
// `ws` must hold workspace_layout(s).total() bytes.
void enqueue_convolution(const ConvShape& s, float* d_out, const float* d_image,
                         const float* d_kernel, Op op, void* ws, hipStream_t stream);

// Building blocks (all asynchronous on `stream`):
//   forward_xy  real volume -> half spectrum S (x-R2C, y-FFT); with `kernel`
//               also the compact kernel's rows Kx and planes Kxy
//   z_pass      S <- IFFTz(FFTz(S) * op(FFTz(operand))) / N, operand the
//               compact kernel planes (dense = false) or a second spectrum
//   inverse_yx  S -> real volume (inverse y-FFT, x-C2R), unnormalised
void forward_xy(const ConvShape& s, const float* src, float2* S, const float* kernel, float2* Kx,
                float2* Kxy, hipStream_t stream, const VPad* vp = nullptr);
// the single passes forward_xy / inverse_yx are made of; with `vp` the x
// passes read a virtually padded image and write its cropped interior
void pass_x_forward(const ConvShape& s, const float* src, float2* S, const float* kernel,
                    float2* Kx, hipStream_t stream, const VPad* vp = nullptr);
void pass_y_forward(const ConvShape& s, float2* S, const float2* Kx, float2* Kxy,
                    hipStream_t stream);
void pass_y_inverse(const ConvShape& s, float2* S, hipStream_t stream);
void pass_x_inverse(const ConvShape& s, const float2* S, float* dst, hipStream_t stream,
                    const VPad* vp = nullptr);
// out_planes / out_shift: write only result planes [0, out_planes), result
// plane p going to (p + out_shift) mod nz (out_planes >= nz: all, in place)
void z_pass(const ConvShape& s, float2* S, const float2* operand, bool dense, bool conj,
            hipStream_t stream, int out_planes = 1 << 30, int out_shift = 0);
void inverse_yx(const ConvShape& s, float2* S, float* dst, hipStream_t stream,
                const VPad* vp = nullptr);

// compute units of the current device (queried once per device)
int cu_count();

// Whole-plane x and y transforms (plane_kernels.hip) for plane extents whose
// half-spectrum plane, or half of it, fits one workgroup's LDS: forward_xy's
// image spectrum and kernel planes (no kernel rows) in one launch,
// inverse_yx in one launch; each direction is chosen on its own.
// Whether direction dir (-1 forward, +1 inverse) runs on plane kernels.
bool plane_path(const ConvShape& s, int dir);
// Whether it runs the split kernels (two workgroups per plane), which read
// the stage table of the half length plane_half_length(s, dir) through
// PlaneArgs::hx_stw / hy_stw.
bool plane_split(const ConvShape& s, int dir);
int plane_half_length(const ConvShape& s, int dir);
void plane_launch(const ConvShape& s, const PlaneArgs& a, int dir, hipStream_t stream);
// Transform lengths whose tables a convolution of shape s reads (nx, ny, nz
// and the split kernels' half lengths, without repeats); returns the count
// (at most 5).
int table_lengths(const ConvShape& s, int* L);
// src null: the kernel's planes only; kernel null: the image's planes only
void plane_forward(const ConvShape& s, const float* src, float2* S, const float* kernel,
                   float2* Kxy, hipStream_t stream);
void plane_inverse(const ConvShape& s, const float2* S, float* dst, hipStream_t stream);

// Boundary-mode convolution (FCAMD_BOUNDARY_*): the circular convolution
// of a virtually padded volume, cropped.  Each axis d with k_d > 1 is padded
// to a fast length P_d >= n_d + k_d - 1 with the image placed at
// lo_d = k_d - 1 - c_d, so no output of the interior wraps around.
struct PaddedProblem {
  ConvShape padded;  // transform shape (P_x, P_y, P_z) with the kernel
  VPad vp;
};
PaddedProblem padded_problem(const ConvShape& s, int boundary);
// smallest fast transform length >= n (a 7-smooth compile-time length when
// one is within 25 %, else the smallest 7-smooth length)
int fast_length(int n);
void enqueue_convolution_padded(const PaddedProblem& p, float* d_out, const float* d_image,
                                const float* d_kernel, Op op, void* ws, hipStream_t stream);

// Shapes with an axis longer than one transform tile (x > 10239, y > 10240,
// z > 6826) take the unfused long-axis path (long_path.cpp): four-step FFTs
// through HBM for such axes, the kernel placed on the image grid.
bool needs_long_path(const ConvShape& s);
// every table the long-axis path reads for shape s, created now (a first use
// while `stream` captures is refused)
void long_prepare_tables(const ConvShape& s, hipStream_t stream);
void validate_long_shape(const ConvShape& s);
size_t long_workspace_bytes(const ConvShape& s);
void enqueue_convolution_long(const ConvShape& s, float* d_out, const float* d_image,
                              const float* d_kernel, Op op, void* ws, hipStream_t stream);
// its two halves, for plans: the kernel's spectrum into the workspace (every
// table the path reads is created first), then the image's forward
// transform, the product with that spectrum and the inverse
void long_kernel_spectrum(const ConvShape& s, const float* d_kernel, void* ws,
                          hipStream_t stream);
void long_apply_spectrum(const ConvShape& s, float* d_out, const float* d_image, Op op, void* ws,
                         hipStream_t stream);
// The PSF gradient (enqueue_grad_kernel) on the long-axis path, in a
// long_workspace_bytes(s) workspace: the cross-correlation of g and the
// image through the unfused passes, then the gather of the support.
void enqueue_grad_kernel_long(const ConvShape& s, float* d_dkernel, const float* d_grad_out,
                              const float* d_image, void* ws, hipStream_t stream);
// exp(-2 pi i j / L), j < L, on the current device (created once; a first
// use while `stream` captures is refused like fft_desc's)
const float2* long_twiddle_table(int L, hipStream_t stream);
// The largest prime factor of L when the one-tile plan of L runs it as a
// direct stage (above 31, with Bluestein's length M past 6826), else 0.
int direct_stage_prime(int L);
// Tables of a length-L transform run as Bluestein's chirp-z over a circular
// convolution of length M = M1 * M2 (M >= 2L - 1, transformed by the
// four-step of that split, or in one tile when M1 = 0): chirp[n] =
// exp(-i pi n^2 / L) for n < L, and bfft the length-M spectrum of the
// conjugate chirp sequence, scaled by 1/M and stored in the forward
// four-step's output order (frequency k1 + M1 k2 at position k1 M2 + k2;
// natural order when M1 = 0).  Created once per device like the twiddle
// tables.
void long_chirp_tables(int L, int M, int M1, hipStream_t stream, const float2** chirp,
                       const float2** bfft);
// dst <- scale * dst * src (conj(src) when conj != 0) over n complex values
void launch_modulate(float2* dst, const float2* src, long n, float scale, int conj,
                     hipStream_t stream);

// Gradient of sum(g * conv(im, k)) with respect to the kernel taps:
// dk[j] = sum_i g[i] im[(i - (j - c)) mod n] on the kernel support.
struct GradWorkspaceLayout {
  size_t spectrum;  // two spectra of the image's size
  size_t corr;      // the correlation on the kz planes the gather reads
  size_t total() const { return 2 * spectrum + corr; }
};
GradWorkspaceLayout grad_workspace_layout(const ConvShape& s);
void enqueue_grad_kernel(const ConvShape& s, float* d_dkernel, const float* d_grad_out,
                         const float* d_image, void* ws, hipStream_t stream);
// dk[jz][jy][jx] = u[(jz-cz) mod nz][(jy-cy) mod ny][(jx-cx) mod nx]
void launch_gather_kernel_support(const ConvShape& s, const float* u, float* dk,
                                  hipStream_t stream);
// The same gather from the y-inverted half-spectrum rows of the kz support
// planes (plane jz holding z-plane (jz - cz) mod nz, row pitch nxp), with the
// x-C2R of each support row evaluated at the kx support positions only; tw
// is the length-nx table exp(-2 pi i j / nx).
void launch_gather_support_c2r(const ConvShape& s, const float2* Y, long nxp, const float2* tw,
                               float* dk, hipStream_t stream);
// its LDS bytes (the row, the table and the partial sums)
size_t gather_support_c2r_lds(int nx);

// Per-device pooled resources for callers that pass no workspace.
struct DevicePool {
  std::mutex mu;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  float* img = nullptr;
  size_t img_bytes = 0;
  float* ker = nullptr;
  size_t ker_bytes = 0;
  float* out = nullptr;  // second volume of the slab pipeline
  size_t out_bytes = 0;
  hipStream_t stream = nullptr;
};

DevicePool& device_pool(int dev);
void* pool_reserve(void** p, size_t* cap, size_t bytes);

// RAII: makes `dev` current and restores the caller's device on exit.
class DeviceGuard {
 public:
  explicit DeviceGuard(int dev);
  ~DeviceGuard();

 private:
  int prev_ = -1;
};

}  // namespace fcamd
