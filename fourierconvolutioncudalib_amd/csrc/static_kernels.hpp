// static_kernels.hpp -- compile-time-length pass kernels (register-resident
// Stockham FFTs) for the lengths registered in static_kernels.hip.  Each
// entry returns false when no compiled kernel covers the length, and the
// driver then uses the runtime-length kernels.
#pragma once

#include "conv_driver.hpp"

namespace fcamd {

// x-R2C of the image rows; with `kernel` non-null the same launch also
// transforms the compact kernel's rows into `krows`.  With `vp` the image
// rows are those of a virtually padded volume (VPad) and the inverse writes
// only its cropped interior.
bool static_xfwd(const ConvShape& s, const FftDesc& fx, const float* src, float2* dst,
                 const float* kernel, float2* krows, hipStream_t st, const VPad* vp = nullptr);
bool static_xinv(const ConvShape& s, const FftDesc& fx, const float2* src, float* dst,
                 hipStream_t st, const VPad* vp = nullptr);
// y-FFT of the spectrum in place; with `krows` non-null the same launch also
// turns the kernel rows into the kernel planes `kplanes`.
bool static_ycol(const ConvShape& s, const FftDesc& fy, float2* S, int dir, const float2* krows,
                 float2* kplanes, hipStream_t st);
// The same y-FFT of the spectrum (no kernel rows) on the plain column
// variant with the largest tile that still gives each of `cus` compute units
// a workgroup, or the smallest tile when none does, whatever the tuner
// selected: for passes over a few planes, which the tuner does not time.
bool static_ycol_fill(const ConvShape& s, const FftDesc& fy, float2* S, int dir, int cus,
                      hipStream_t st);
bool static_zfused(const ConvShape& s, const ZArgs& a, hipStream_t st);
void static_kernels_allow_large_lds();

}  // namespace fcamd
