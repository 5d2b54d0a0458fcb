// kernels_long.hpp -- argument block and launchers of the long-axis path
// (kernels_long.hip).
#pragma once

#include "fc_internal.hpp"

namespace fcamd {

// Column FFTs of length fft.L over a [outer][L][cols] view.  Outer row q
// starts at (q / outer_div) * hi + (q % outer_div) * lo (outer_div 0: q * hi)
// in src and dst (which may alias); position p and column c add
// p * elem + c * col.  twiddle: 0 none, 1 forward four-step (after the FFT),
// 2 inverse four-step (before it), with exponent position * n2 over tw_len,
// n2 the column index (tw_by_col) or q mod outer_div.
struct LColArgs {
  FftDesc fft;
  int tx;       // columns per tile
  long ncols;   // columns per outer row
  long nouter;  // outer rows
  const float2* src;
  float2* dst;
  long src_hi, src_lo, src_elem, src_col;
  long dst_hi, dst_lo, dst_elem, dst_col;
  long outer_div;
  int twiddle;
  int tw_by_col;
  const float2* tw_long;  // exp(-2 pi i j / tw_len), j < tw_len
  long tw_len;
};

void launch_lcol(const LColArgs& a, int dir, hipStream_t st);
// real rows <-> packed complex rows (two rows each) and their spectra; L1,
// L2: the four-step split of nx (L1 = 0: natural frequency order)
void launch_xpack(const float* v, float2* z, long nrows, int nx, hipStream_t st);
void launch_xsep(const float2* z, float2* S, long nrows, int nx, long srow, int L1, int L2,
                 hipStream_t st);
void launch_xunpack(const float2* S, float2* z, long nrows, int nx, long srow, int L1, int L2,
                    hipStream_t st);
void launch_xreal(const float2* z, float* v, long nrows, int nx, hipStream_t st);
// Bluestein's chirp-z steps for an axis of length L with no split into two
// tiles (kernels_long.hip): lines x(o, p, c) at o * hi + p * elem + c,
// scratch W[nouter][M][ncols]; dir -1 forward, +1 inverse
void launch_chirp_pre(const float2* x, long hi, long elem, long ncols, long nouter, int L,
                      float2* W, int M, const float2* chirp, int dir, hipStream_t st);
void launch_chirp_mul(float2* W, long ncols, long nouter, int M, const float2* bfft, int dir,
                      hipStream_t st);
void launch_chirp_post(const float2* W, float2* x, long hi, long elem, long ncols, long nouter,
                       int L, int M, const float2* chirp, int dir, hipStream_t st);
void launch_fill_zero(float* v, long n, hipStream_t st);
void launch_scatter_kernel(const float* k, float* v, int nx, int ny, int nz, int kx, int ky,
                           int kz, int cx, int cy, int cz, hipStream_t st);
void long_kernels_allow_large_lds();

}  // namespace fcamd
