// static_kernels.hpp -- compile-time-length pass kernels (register-resident
// Stockham FFTs) for the lengths registered in static_kernels.hip.  Each
// entry returns false when no compiled kernel covers the length, and the
// driver then uses the runtime-length kernels.
#pragma once

#include "conv_driver.hpp"

namespace fcamd {

bool static_xfwd(const ConvShape& s, const FftDesc& fx, const float* src, float2* dst,
                 hipStream_t st);
bool static_xinv(const ConvShape& s, const FftDesc& fx, const float2* src, float* dst,
                 hipStream_t st);
bool static_ycol(const ConvShape& s, const FftDesc& fy, float2* S, int dir, hipStream_t st);
bool static_zfused(const ConvShape& s, const ZArgs& a, hipStream_t st);
void static_kernels_allow_large_lds();

}  // namespace fcamd
