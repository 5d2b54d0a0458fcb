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
// inverse undoes.  A length with no split into two tiles (a prime factor
// above one tile), or one whose one-tile plan would run a direct stage of a
// prime above kChirpPrime, runs as Bluestein's chirp-z: a circular
// convolution of a 7-smooth length M >= 2L - 1 (in one tile, or split into
// two), through a scratch of M positions per line; its result is in natural
// frequency order.
#include <algorithm>
#include <cmath>
#include <string>

#include "conv_driver.hpp"
#include "kernels_long.hpp"

namespace fcamd {
namespace {

// longest complex length the runtime-length engine holds in one tile (two
// LDS buffers of 160 KiB)
constexpr int kTileLen = 10240;

// A direct p-point stage costs p complex multiply-adds per output.  On
// MI355X, x axes of L x 64 x 8 volumes (5^3 kernel) on their direct stages
// took 32.7 ms at 4099 and 206 ms at 9973 (0.47 and 1.15 ms on the
// chirp-z), 0.67-1.14 ms at 3431 = 47 * 73, 3599 = 59 * 61, 3737 = 37 * 101
// and 4661 = 59 * 79 (about 0.45 ms on the chirp-z at these lengths), and
// 0.30 ms at 3700 = 37 * 100.  Axes whose one-tile plan would run a direct
// stage of a prime above this take the chirp-z.
constexpr int kChirpPrime = 41;

bool chirp_length(int L) { return direct_stage_prime(L) > kChirpPrime; }

struct AxisSplit {
  int L, L1, L2;      // four-step L = L1 * L2; L1 = 0 and M = 0: one tile
  int M = 0, M1 = 0;  // chirp-z over the four-step M = M1 * (M / M1)
  bool one_tile() const { return L1 == 0 && M == 0; }
};

// the divisor nearest sqrt(L) leaving both parts within a tile, or 0
int two_tile_divisor(long L, int tile) {
  for (long d = (long)std::sqrt((double)L); d >= 2; --d)
    if (L % d == 0 && L / d <= tile && d <= tile) return (int)d;
  return 0;
}

bool seven_smooth(long m) {
  for (int f : {2, 3, 5, 7})
    while (m % f == 0) m /= f;
  return m == 1;
}

// one tile when L <= tile (unless `force`), else L1 * L2 with both in a
// tile, else the chirp-z over the smallest 7-smooth M >= 2L - 1 in one tile
// or split into two; lengths with a large direct prime stage go to the
// chirp-z straight away
AxisSplit split_axis(int L, int tile, bool force = false) {
  if (!chirp_length(L)) {
    if (L <= tile && !force) return {L, 0, 0};
    if (const int d = two_tile_divisor(L, tile)) return {L, d, L / d};
  }
  for (long m = 2L * L - 1; m <= (long)tile * tile; ++m)
    if (seven_smooth(m)) {
      if (m <= tile) return {L, 0, 0, (int)m, 0};
      if (const int d = two_tile_divisor(m, tile)) return {L, 0, 0, (int)m, d};
    }
  throw Error{kUnsupportedSize,
              "transform length " + std::to_string(L) +
                  " exceeds one tile and has no split into two tiles, nor a chirp-z length "
                  "that has one"};
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

void chirp_z(const AxisSplit& sp, long ncols, long nouter, float2* data, long hi, long elem,
             int dir, float2* W, hipStream_t st);

// In-place FFT (dir -1 forward, +1 inverse) along an axis of length L with
// position stride `elem`, over `ncols` contiguous columns per outer row and
// `nouter` outer rows of stride `hi`; W: the chirp-z scratch, ncols * nouter
// * sp.M positions.
void axis_fft(const AxisSplit& sp, long ncols, long nouter, float2* data, long hi, long elem,
              int dir, float2* W, hipStream_t st) {
  if (sp.M > 0) {
    chirp_z(sp, ncols, nouter, data, hi, elem, dir, W, st);
    return;
  }
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
void row_fft(const AxisSplit& sp, long npairs, float2* z, int dir, float2* W, hipStream_t st) {
  const int nx = sp.L;
  if (sp.M > 0) {
    chirp_z(sp, 1, npairs, z, nx, 1, dir, W, st);
    return;
  }
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

// Bluestein's chirp-z along the axis (kernels_long.hip: launch_chirp_*); the
// length-M transforms are four-steps, over contiguous rows when ncols = 1
void chirp_z(const AxisSplit& sp, long ncols, long nouter, float2* data, long hi, long elem,
             int dir, float2* W, hipStream_t st) {
  FC_REQUIRE(W != nullptr, kInternal, "chirp-z without its scratch");
  const float2* chirp = nullptr;
  const float2* bfft = nullptr;
  long_chirp_tables(sp.L, sp.M, sp.M1, st, &chirp, &bfft);
  const AxisSplit m{sp.M, sp.M1, sp.M1 ? sp.M / sp.M1 : 0};
  launch_chirp_pre(data, hi, elem, ncols, nouter, sp.L, W, sp.M, chirp, dir, st);
  if (ncols == 1)
    row_fft(m, nouter, W, -1, nullptr, st);
  else
    axis_fft(m, ncols, nouter, W, (long)sp.M * ncols, ncols, -1, nullptr, st);
  launch_chirp_mul(W, ncols, nouter, sp.M, bfft, dir, st);
  if (ncols == 1)
    row_fft(m, nouter, W, +1, nullptr, st);
  else
    axis_fft(m, ncols, nouter, W, (long)sp.M * ncols, ncols, +1, nullptr, st);
  launch_chirp_post(W, data, hi, elem, ncols, nouter, sp.L, sp.M, chirp, dir, st);
}

AxisSplit x_split(const ConvShape& s) {
  // the single-tile x pass holds nx + 1 complex values per row pair; past
  // that the packed rows always take the four-step (whose parts are
  // column FFTs of at most kTileLen) or the chirp-z
  return s.nx <= kTileLen - 1 && !chirp_length(s.nx) ? AxisSplit{s.nx, 0, 0}
                                                      : split_axis(s.nx, kTileLen, true);
}

struct LongLayout {
  size_t spectrum, kspectrum, kvol, packed, chirpz;
  size_t total() const { return spectrum + kspectrum + kvol + packed + chirpz; }
};

LongLayout long_layout(const ConvShape& s) {
  const size_t nxp = (size_t)padded_bins(s.nx);
  const size_t rows = (size_t)s.ny * s.nz;
  LongLayout w{};
  w.spectrum = align256(sizeof(float2) * nxp * rows);
  w.kspectrum = w.spectrum;
  w.kvol = align256(sizeof(float) * (size_t)s.voxels());
  const AxisSplit sx = x_split(s), sy = split_axis(s.ny, kTileLen), sz = split_axis(s.nz, kTileLen);
  w.packed = sx.one_tile() ? 0 : align256(sizeof(float2) * ((rows + 1) / 2) * s.nx);
  // one chirp-z scratch for the largest axis that needs one
  const size_t nb = (size_t)half_bins(s.nx);
  size_t cz = 0;
  if (sx.M > 0) cz = std::max(cz, (rows + 1) / 2 * (size_t)sx.M);
  if (s.ny > 1 && sy.M > 0) cz = std::max(cz, nb * s.nz * (size_t)sy.M);
  if (s.nz > 1 && sz.M > 0) cz = std::max(cz, nb * s.ny * (size_t)sz.M);
  w.chirpz = align256(sizeof(float2) * cz);
  return w;
}

// W: the chirp-z scratch (LongLayout::chirpz)
void forward(const ConvShape& s, const float* v, float2* S, float2* Z, float2* W,
             hipStream_t st) {
  const long nxp = padded_bins(s.nx), rows = (long)s.ny * s.nz, plane = (long)s.ny * nxp;
  const int nb = half_bins(s.nx);
  const AxisSplit sx = x_split(s);
  if (sx.one_tile()) {
    pass_x_forward(s, v, S, nullptr, nullptr, st);
  } else {
    launch_xpack(v, Z, rows, s.nx, st);
    row_fft(sx, (rows + 1) / 2, Z, -1, W, st);
    launch_xsep(Z, S, rows, s.nx, nxp, sx.L1, sx.L2, st);
  }
  if (s.ny > 1) axis_fft(split_axis(s.ny, kTileLen), nb, s.nz, S, plane, nxp, -1, W, st);
  if (s.nz > 1) axis_fft(split_axis(s.nz, kTileLen), nb, s.ny, S, nxp, plane, -1, W, st);
}

void inverse(const ConvShape& s, float2* S, float* v, float2* Z, float2* W, hipStream_t st) {
  const long nxp = padded_bins(s.nx), rows = (long)s.ny * s.nz, plane = (long)s.ny * nxp;
  const int nb = half_bins(s.nx);
  if (s.nz > 1) axis_fft(split_axis(s.nz, kTileLen), nb, s.ny, S, nxp, plane, +1, W, st);
  if (s.ny > 1) axis_fft(split_axis(s.ny, kTileLen), nb, s.nz, S, plane, nxp, +1, W, st);
  const AxisSplit sx = x_split(s);
  if (sx.one_tile()) {
    pass_x_inverse(s, S, v, st);
  } else {
    launch_xunpack(S, Z, rows, s.nx, nxp, sx.L1, sx.L2, st);
    row_fft(sx, (rows + 1) / 2, Z, +1, W, st);
    launch_xreal(Z, v, rows, s.nx, st);
  }
}

// Creates every table the path reads before its first launch, so that a
// first use inside a stream capture is refused with nothing recorded.
void prepare_tables(const ConvShape& s, hipStream_t st) {
  auto split_tables = [&](int L, int L1) {
    if (L1 > 0) {
      (void)fft_desc(L1, st);
      (void)fft_desc(L / L1, st);
      (void)long_twiddle_table(L, st);
    } else {
      (void)fft_desc(L, st);
    }
  };
  auto axis_tables = [&](const AxisSplit& a) {
    if (a.M > 0) {
      const float2* c = nullptr;
      const float2* b = nullptr;
      long_chirp_tables(a.L, a.M, a.M1, st, &c, &b);
      split_tables(a.M, a.M1);
    } else {
      split_tables(a.L, a.L1);
    }
  };
  axis_tables(x_split(s));
  if (s.ny > 1) axis_tables(split_axis(s.ny, kTileLen));
  if (s.nz > 1) axis_tables(split_axis(s.nz, kTileLen));
}

}  // namespace

void long_prepare_tables(const ConvShape& s, hipStream_t st) {
  validate_long_shape(s);
  prepare_tables(s, st);
}

bool needs_long_path(const ConvShape& s) {
  return s.nx > kTileLen - 1 || s.ny > kTileLen || s.nz > 6826 || s.ny > 65535 ||
         s.nz > 65535 || chirp_length(s.nx) || chirp_length(s.ny) || chirp_length(s.nz);
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

namespace {
struct LongRegions {
  float2 *S, *K;
  float* kv;
  float2 *Z, *W;
};
LongRegions long_regions(const ConvShape& s, void* ws) {
  const LongLayout w = long_layout(s);
  char* base = static_cast<char*>(ws);
  LongRegions r{};
  r.S = reinterpret_cast<float2*>(base);
  r.K = reinterpret_cast<float2*>(base + w.spectrum);
  r.kv = reinterpret_cast<float*>(base + w.spectrum + w.kspectrum);
  r.Z = reinterpret_cast<float2*>(base + w.spectrum + w.kspectrum + w.kvol);
  r.W = w.chirpz ? reinterpret_cast<float2*>(base + w.spectrum + w.kspectrum + w.kvol + w.packed)
                 : nullptr;
  return r;
}
}  // namespace

void long_kernel_spectrum(const ConvShape& s, const float* d_kernel, void* ws, hipStream_t st) {
  validate_long_shape(s);
  driver_init();
  prepare_tables(s, st);
  const LongRegions r = long_regions(s, ws);
  // the kernel on the image grid, then its spectrum
  launch_fill_zero(r.kv, s.voxels(), st);
  launch_scatter_kernel(d_kernel, r.kv, s.nx, s.ny, s.nz, s.kx, s.ky, s.kz, s.cx, s.cy, s.cz, st);
  forward(s, r.kv, r.K, r.Z, r.W, st);
  FC_HIP_CHECK(hipGetLastError());
}

void long_apply_spectrum(const ConvShape& s, float* d_out, const float* d_image, Op op, void* ws,
                         hipStream_t st) {
  const LongRegions r = long_regions(s, ws);
  forward(s, d_image, r.S, r.Z, r.W, st);
  const long n = (long)padded_bins(s.nx) * s.ny * s.nz;
  launch_modulate(r.S, r.K, n, (float)(1.0 / (double)s.voxels()), op == Op::kCorrelate ? 1 : 0,
                  st);
  inverse(s, r.S, d_out, r.Z, r.W, st);
  FC_HIP_CHECK(hipGetLastError());
}

void enqueue_convolution_long(const ConvShape& s, float* d_out, const float* d_image,
                              const float* d_kernel, Op op, void* ws, hipStream_t st) {
  long_kernel_spectrum(s, d_kernel, ws, st);
  long_apply_spectrum(s, d_out, d_image, op, ws, st);
}

void enqueue_grad_kernel_long(const ConvShape& s, float* d_dkernel, const float* d_grad_out,
                              const float* d_image, void* ws, hipStream_t st) {
  validate_long_shape(s);
  driver_init();
  prepare_tables(s, st);
  const LongRegions r = long_regions(s, ws);
  // u = IFFT(FFT(g) conj(FFT(im))) / N, the cross-correlation, into the
  // kernel-volume region; dk[j] = u[(j - c) mod n] on the support
  forward(s, d_grad_out, r.S, r.Z, r.W, st);
  forward(s, d_image, r.K, r.Z, r.W, st);
  const long n = (long)padded_bins(s.nx) * s.ny * s.nz;
  launch_modulate(r.S, r.K, n, (float)(1.0 / (double)s.voxels()), 1, st);
  inverse(s, r.S, r.kv, r.Z, r.W, st);
  launch_gather_kernel_support(s, r.kv, d_dkernel, st);
  FC_HIP_CHECK(hipGetLastError());
}

}  // namespace fcamd
