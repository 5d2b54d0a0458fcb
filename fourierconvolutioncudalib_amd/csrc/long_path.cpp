// long_path.cpp -- convolution of shapes with an axis longer than one LDS
// tile (the single-tile passes take x <= 10239, y <= 10240, z <= 6826).
//
// cuFFT plans any size (src/convolution3Dfft.cu:519,544); here an axis that
// does not fit one tile is split L = L1 * L2 (the divisor nearest sqrt(L),
// both parts within a tile) and transformed in four steps through HBM
// (kernels_long.hip).  The path is unfused, the reference's own structure:
//   forward(image) -> S; the kernel placed circularly on the image grid
//   (fftShiftKernel, src/convolution3Dfft.cu:128-166) -> forward -> K;
//   S <- S * K / N (modulateAndNormalize_kernel, :47-62); inverse(S) -> out.
// forward = x-R2C (two real rows per complex row), y-FFT, z-FFT; inverse the
// reverse.  A four-step axis leaves its frequencies in a transposed order,
// the same for S and K, which the pointwise product does not see and the
// inverse undoes.
#include <cmath>
#include <string>

#include "conv_driver.hpp"
#include "kernels_long.hpp"

namespace fcamd {
namespace {

// longest complex length the runtime-length engine holds in one tile (two
// LDS buffers of 160 KiB)
constexpr int kTileLen = 10240;

struct AxisSplit {
  int L, L1, L2;  // L1 = 0: one tile
};

// one tile when L <= tile (unless `force`), else L1 * L2 with both in a tile
AxisSplit split_axis(int L, int tile, bool force = false) {
  if (L <= tile && !force) return {L, 0, 0};
  for (int d = (int)std::sqrt((double)L); d >= 2; --d)
    if (L % d == 0 && L / d <= tile && d <= tile) return {L, d, L / d};
  throw Error{kUnsupportedSize, "transform length " + std::to_string(L) +
                                    " exceeds one tile and has no split into two tiles"};
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// columns per tile: up to 16, within 64 KiB of LDS for the two buffers
int lcol_tile(int L, long ncols) {
  int tx = 16;
  while (tx > 1 && (2 * sizeof(float2) * (size_t)L * tx > 65536 || tx > ncols)) tx /= 2;
  return tx;
}

LColArgs lcol(int L, long ncols, long nouter, float2* data, long hi, long lo, long div, long elem,
              hipStream_t st) {
  LColArgs a{};
  a.fft = fft_desc(L, st);
  a.tx = lcol_tile(a.fft.buf(), ncols);
  a.ncols = ncols;
  a.nouter = nouter;
  a.src = data;
  a.dst = data;
  a.src_hi = a.dst_hi = hi;
  a.src_lo = a.dst_lo = lo;
  a.outer_div = div;
  a.src_elem = a.dst_elem = elem;
  a.src_col = a.dst_col = 1;
  return a;
}

// In-place FFT (dir -1 forward, +1 inverse) along an axis of length L with
// position stride `elem`, over `ncols` contiguous columns per outer row and
// `nouter` outer rows of stride `hi`.
void axis_fft(const AxisSplit& sp, long ncols, long nouter, float2* data, long hi, long elem,
              int dir, hipStream_t st) {
  if (sp.L1 == 0) {
    launch_lcol(lcol(sp.L, ncols, nouter, data, hi, 0, 0, elem, st), dir, st);
    return;
  }
  // element n1*L2 + n2 along the axis
  LColArgs A = lcol(sp.L1, ncols, nouter * sp.L2, data, hi, elem, sp.L2, elem * sp.L2, st);
  A.tw_long = long_twiddle_table(sp.L, st);
  A.tw_len = sp.L;
  LColArgs B = lcol(sp.L2, ncols, nouter * sp.L1, data, hi, elem * sp.L2, sp.L1, elem, st);
  if (dir < 0) {
    A.twiddle = 1;
    launch_lcol(A, -1, st);
    launch_lcol(B, -1, st);
  } else {
    A.twiddle = 2;
    launch_lcol(B, +1, st);
    launch_lcol(A, +1, st);
  }
}

// Complex FFTs of the packed rows z[npairs][nx] (contiguous rows).
void row_fft(const AxisSplit& sp, long npairs, float2* z, int dir, hipStream_t st) {
  const int nx = sp.L;
  if (sp.L1 == 0) {
    launch_lcol(lcol(nx, 1, npairs, z, nx, 0, 0, 1, st), dir, st);
    return;
  }
  // step A over n1 (stride L2) for the L2 contiguous columns n2, twiddle by
  // column; step B over n2 within each contiguous block of L2
  LColArgs A = lcol(sp.L1, sp.L2, npairs, z, nx, 0, 0, sp.L2, st);
  A.tw_long = long_twiddle_table(nx, st);
  A.tw_len = nx;
  A.tw_by_col = 1;
  LColArgs B = lcol(sp.L2, 1, npairs * sp.L1, z, sp.L2, 0, 0, 1, st);
  if (dir < 0) {
    A.twiddle = 1;
    launch_lcol(A, -1, st);
    launch_lcol(B, -1, st);
  } else {
    A.twiddle = 2;
    launch_lcol(B, +1, st);
    launch_lcol(A, +1, st);
  }
}

struct LongLayout {
  size_t spectrum, kspectrum, kvol, packed;
  size_t total() const { return spectrum + kspectrum + kvol + packed; }
};

LongLayout long_layout(const ConvShape& s) {
  const size_t nxp = (size_t)padded_bins(s.nx);
  const size_t rows = (size_t)s.ny * s.nz;
  LongLayout w{};
  w.spectrum = align256(sizeof(float2) * nxp * rows);
  w.kspectrum = w.spectrum;
  w.kvol = align256(sizeof(float) * (size_t)s.voxels());
  w.packed = s.nx > kTileLen - 1 ? align256(sizeof(float2) * ((rows + 1) / 2) * s.nx) : 0;
  return w;
}

AxisSplit x_split(const ConvShape& s) {
  // the single-tile x pass holds nx + 1 complex values per row pair; past
  // that the packed rows always take the four-step (whose parts are
  // column FFTs of at most kTileLen)
  return s.nx <= kTileLen - 1 ? AxisSplit{s.nx, 0, 0} : split_axis(s.nx, kTileLen, true);
}

void forward(const ConvShape& s, const float* v, float2* S, float2* Z, hipStream_t st) {
  const long nxp = padded_bins(s.nx), rows = (long)s.ny * s.nz, plane = (long)s.ny * nxp;
  const int nb = half_bins(s.nx);
  const AxisSplit sx = x_split(s);
  if (sx.L1 == 0) {
    pass_x_forward(s, v, S, nullptr, nullptr, st);
  } else {
    launch_xpack(v, Z, rows, s.nx, st);
    row_fft(sx, (rows + 1) / 2, Z, -1, st);
    launch_xsep(Z, S, rows, s.nx, nxp, sx.L1, sx.L2, st);
  }
  if (s.ny > 1) axis_fft(split_axis(s.ny, kTileLen), nb, s.nz, S, plane, nxp, -1, st);
  if (s.nz > 1) axis_fft(split_axis(s.nz, kTileLen), nb, s.ny, S, nxp, plane, -1, st);
}

void inverse(const ConvShape& s, float2* S, float* v, float2* Z, hipStream_t st) {
  const long nxp = padded_bins(s.nx), rows = (long)s.ny * s.nz, plane = (long)s.ny * nxp;
  const int nb = half_bins(s.nx);
  if (s.nz > 1) axis_fft(split_axis(s.nz, kTileLen), nb, s.ny, S, nxp, plane, +1, st);
  if (s.ny > 1) axis_fft(split_axis(s.ny, kTileLen), nb, s.nz, S, plane, nxp, +1, st);
  const AxisSplit sx = x_split(s);
  if (sx.L1 == 0) {
    pass_x_inverse(s, S, v, st);
  } else {
    launch_xunpack(S, Z, rows, s.nx, nxp, sx.L1, sx.L2, st);
    row_fft(sx, (rows + 1) / 2, Z, +1, st);
    launch_xreal(Z, v, rows, s.nx, st);
  }
}

}  // namespace

bool needs_long_path(const ConvShape& s) {
  return s.nx > kTileLen - 1 || s.ny > kTileLen || s.nz > 6826 || s.ny > 65535 || s.nz > 65535;
}

void validate_long_shape(const ConvShape& s) {
  FC_REQUIRE(s.nx >= 1 && s.ny >= 1 && s.nz >= 1, kInvalidArgument,
             "image dimensions must be >= 1");
  FC_REQUIRE(s.kx >= 1 && s.ky >= 1 && s.kz >= 1, kInvalidArgument,
             "kernel dimensions must be >= 1");
  FC_REQUIRE(s.kx <= s.nx && s.ky <= s.ny && s.kz <= s.nz, kInvalidArgument,
             "kernel extent exceeds image extent on some axis");
  FC_REQUIRE(s.cx >= 0 && s.cx < s.kx && s.cy >= 0 && s.cy < s.ky && s.cz >= 0 && s.cz < s.kz,
             kInvalidArgument, "kernel centre outside the kernel");
  (void)x_split(s);
  (void)split_axis(s.ny, kTileLen);
  (void)split_axis(s.nz, kTileLen);
}

size_t long_workspace_bytes(const ConvShape& s) { return long_layout(s).total(); }

void enqueue_convolution_long(const ConvShape& s, float* d_out, const float* d_image,
                              const float* d_kernel, Op op, void* ws, hipStream_t st) {
  validate_long_shape(s);
  driver_init();
  const LongLayout w = long_layout(s);
  char* base = static_cast<char*>(ws);
  float2* S = reinterpret_cast<float2*>(base);
  float2* K = reinterpret_cast<float2*>(base + w.spectrum);
  float* kv = reinterpret_cast<float*>(base + w.spectrum + w.kspectrum);
  float2* Z = reinterpret_cast<float2*>(base + w.spectrum + w.kspectrum + w.kvol);
  // the kernel on the image grid, then both spectra
  launch_fill_zero(kv, s.voxels(), st);
  launch_scatter_kernel(d_kernel, kv, s.nx, s.ny, s.nz, s.kx, s.ky, s.kz, s.cx, s.cy, s.cz, st);
  forward(s, kv, K, Z, st);
  forward(s, d_image, S, Z, st);
  const long n = (long)padded_bins(s.nx) * s.ny * s.nz;
  launch_modulate(S, K, n, (float)(1.0 / (double)s.voxels()), op == Op::kCorrelate ? 1 : 0, st);
  inverse(s, S, d_out, Z, st);
  FC_HIP_CHECK(hipGetLastError());
}

}  // namespace fcamd
