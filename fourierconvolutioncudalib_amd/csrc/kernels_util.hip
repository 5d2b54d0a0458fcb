// kernels_util.hip -- small stand-alone kernels: the unfused pointwise
// spectrum product (src/convolution3Dfft.cu:41-62, modulateAndNormalize) and
// a streaming copy used to calibrate HBM byte counters.
#include "fc_internal.hpp"
#include "fft_radix.hpp"
#include "../../include/fourierconvolution_amd.h"

namespace fcamd {

// dst[i] <- scale * dst[i] * src[i] (conj(src[i]) when conj != 0)
__global__ void __launch_bounds__(256)
modulate_and_normalize(float2* __restrict__ dst, const float2* __restrict__ src, long n,
                       float scale, int conj) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const float2 a = dst[i], b = src[i];
    dst[i] = cscale(conj ? cmulc(a, b) : cmul(a, b), scale);
  }
}

// 16-byte streaming copy
__global__ void __launch_bounds__(256)
stream_copy(float4* __restrict__ dst, const float4* __restrict__ src, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// dk[jz][jy][jx] = u[(jz-cz) mod nz][(jy-cy) mod ny][(jx-cx) mod nx]
__global__ void __launch_bounds__(256)
gather_kernel_support(const float* __restrict__ u, float* __restrict__ dk, int nx, int ny, int nz,
                      int kx, int ky, int kz, int cx, int cy, int cz) {
  const long K = (long)kx * ky * kz;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < K; i += (long)gridDim.x * blockDim.x) {
    const int jx = (int)(i % kx);
    const int jy = (int)((i / kx) % ky);
    const int jz = (int)(i / ((long)kx * ky));
    int x = jx - cx, y = jy - cy, z = jz - cz;
    if (x < 0) x += nx;
    if (y < 0) y += ny;
    if (z < 0) z += nz;
    dk[i] = u[((long)z * ny + y) * nx + x];
  }
}

// dk[jz][jy][jx] = u[jz][(jy-cy) mod ny][(jx-cx) mod nx] straight from the
// support planes' y-inverted half-spectrum rows Y (plane jz of Y holds
// z-plane (jz - cz) mod nz, rotated there by the z pass): one workgroup per
// support row (jz, jy), each output the row's C2R at one x,
//   u(x) = Re Y[0] + sum_{0<k<nx/2} 2 Re(Y[k] e^{2 pi i k x / nx})
//          + Re Y[nx/2] (-1)^x   (even nx),
// the same Hermitian extension the x-C2R pass transforms.  The row and the
// table tw[j] = exp(-2 pi i j / nx) are staged in LDS; each output's sum is
// split over P = blockDim / kx threads and reduced through LDS.
__global__ void __launch_bounds__(256)
gather_support_c2r(const float2* __restrict__ Y, long nxp, const float2* __restrict__ tw,
                   float* __restrict__ dk, int nx, int ny, int kx, int ky, int cx, int cy) {
  extern __shared__ float2 smem[];
  const int nb = nx / 2 + 1;
  float2* row = smem;
  float2* twl = smem + nb;
  float* part = reinterpret_cast<float*>(twl + nx);  // blockDim.x partial sums
  const int jz = blockIdx.x / ky, jy = blockIdx.x - jz * ky;
  int y = jy - cy;
  if (y < 0) y += ny;
  const float2* src = Y + ((long)jz * ny + y) * nxp;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) row[k] = src[k];
  for (int k = threadIdx.x; k < nx; k += blockDim.x) twl[k] = tw[k];
  __syncthreads();
  const bool even = (nx & 1) == 0;
  const int kend = even ? nx / 2 : nb;  // paired bins [1, kend)
  const int P = (int)blockDim.x / kx >= 1 ? (int)blockDim.x / kx : 1;
  const int per = (int)blockDim.x / P;  // outputs per round
  const int C = (kend - 1 + P - 1) / P;  // bins per partial sum
  const int slot = threadIdx.x / P, p = threadIdx.x - slot * P;
  for (int j0 = 0; j0 < kx; j0 += per) {  // workgroup-uniform rounds
    const int jx = j0 + slot;
    const bool act = slot < per && jx < kx;
    int x = 0;
    float pair = 0.f;
    if (act) {
      x = jx - cx;
      if (x < 0) x += nx;
      const int k0 = 1 + p * C, k1 = k0 + C < kend ? k0 + C : kend;
      int e = (int)(((long)k0 * x) % nx);  // (k x) mod nx
      for (int k = k0; k < k1; ++k) {
        // Re(Y e^{+i theta}) with e^{+i theta} = conj(tw[e])
        const float2 h = row[k], w = twl[e];
        pair += h.x * w.x + h.y * w.y;
        e += x;
        if (e >= nx) e -= nx;
      }
    }
    part[threadIdx.x] = pair;
    __syncthreads();
    if (act && p == 0) {
      float sum = 0.f;
      for (int q = 0; q < P; ++q) sum += part[threadIdx.x + q];
      float acc = row[0].x + 2.f * sum;
      if (even) acc += (x & 1) ? -row[nx / 2].x : row[nx / 2].x;
      dk[((long)jz * ky + jy) * kx + jx] = acc;
    }
    __syncthreads();  // the partials are reused by the next round
  }
}

// Richardson-Lucy elementwise steps (unfused iteration and initialisation)
__global__ void __launch_bounds__(256)
rl_mean_views(float* __restrict__ psi, const float* const* __restrict__ views, int nviews, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int v = 0; v < nviews; ++v) s += views[v][i];
    psi[i] = s / (float)nviews;
  }
}

// ratio = phi / max(blur, eps)
__global__ void __launch_bounds__(256)
rl_ratio(float* __restrict__ ratio, const float* __restrict__ phi, const float* __restrict__ blur,
         float eps, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    ratio[i] = phi[i] / fmaxf(blur[i], eps);
}

// acc = (first ? 0 : acc) + back
__global__ void __launch_bounds__(256)
rl_accumulate(float* __restrict__ acc, const float* __restrict__ back, int first, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    acc[i] = first ? back[i] : acc[i] + back[i];
}

// psi = psi * acc * inv_views
__global__ void __launch_bounds__(256)
rl_update(float* __restrict__ psi, const float* __restrict__ acc, float inv_views, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    psi[i] = psi[i] * acc[i] * inv_views;
}

static unsigned ew_blocks(long n) {
  long b = (n + 255) / 256;
  return (unsigned)(b > 8192 ? 8192 : (b < 1 ? 1 : b));
}

void launch_rl_mean(float* psi, const float* const* d_views, int nviews, long n, hipStream_t st) {
  hipLaunchKernelGGL(rl_mean_views, dim3(ew_blocks(n)), dim3(256), 0, st, psi, d_views, nviews, n);
}
void launch_rl_ratio(float* ratio, const float* phi, const float* blur, float eps, long n,
                     hipStream_t st) {
  hipLaunchKernelGGL(rl_ratio, dim3(ew_blocks(n)), dim3(256), 0, st, ratio, phi, blur, eps, n);
}
void launch_rl_accumulate(float* acc, const float* back, int first, long n, hipStream_t st) {
  hipLaunchKernelGGL(rl_accumulate, dim3(ew_blocks(n)), dim3(256), 0, st, acc, back, first, n);
}
void launch_rl_update(float* psi, const float* acc, float inv_views, long n, hipStream_t st) {
  hipLaunchKernelGGL(rl_update, dim3(ew_blocks(n)), dim3(256), 0, st, psi, acc, inv_views, n);
}

}  // namespace fcamd

#include "conv_driver.hpp"

namespace fcamd {
void launch_modulate(float2* dst, const float2* src, long n, float scale, int conj,
                     hipStream_t st) {
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) return;
  hipLaunchKernelGGL(modulate_and_normalize, dim3((unsigned)blocks), dim3(256), 0, st, dst, src, n,
                     scale, conj);
}

size_t gather_support_c2r_lds(int nx) {
  return sizeof(float2) * ((size_t)(nx / 2 + 1) + (size_t)nx) + sizeof(float) * 256;
}

void launch_gather_support_c2r(const ConvShape& s, const float2* Y, long nxp, const float2* tw,
                               float* dk, hipStream_t st) {
  hipLaunchKernelGGL(gather_support_c2r, dim3((unsigned)(s.kz * s.ky)), dim3(256),
                     gather_support_c2r_lds(s.nx), st, Y, nxp, tw, dk, s.nx, s.ny, s.kx, s.ky,
                     s.cx, s.cy);
}

void launch_gather_kernel_support(const ConvShape& s, const float* u, float* dk, hipStream_t st) {
  const long K = (long)s.kx * s.ky * s.kz;
  long blocks = (K + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(gather_kernel_support, dim3((unsigned)blocks), dim3(256), 0, st, u, dk, s.nx,
                     s.ny, s.nz, s.kx, s.ky, s.kz, s.cx, s.cy, s.cz);
}
}  // namespace fcamd

using namespace fcamd;

extern "C" {

// Pointwise complex product on device arrays of n complex values.
FourierConvolutionCUDALib_EXPORT int fcamd_modulate_device(void* d_dst, const void* d_src,
                                                          long long n, float scale, int conj,
                                                          void* stream) {
  if (!d_dst || !d_src || n < 0) return kInvalidArgument;
  if (n == 0) return kOk;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(modulate_and_normalize, dim3((unsigned)blocks), dim3(256), 0,
                     static_cast<hipStream_t>(stream), static_cast<float2*>(d_dst),
                     static_cast<const float2*>(d_src), (long)n, scale, conj);
  return hipGetLastError() == hipSuccess ? kOk : kHipError;
}

// Copies `bytes` (multiple of 16) between device buffers with a plain kernel.
FourierConvolutionCUDALib_EXPORT int fcamd_stream_copy_device(void* d_dst, const void* d_src,
                                                             long long bytes, void* stream) {
  if (!d_dst || !d_src || bytes < 0 || (bytes % 16) != 0) return kInvalidArgument;
  const long n = (long)(bytes / 16);
  if (n == 0) return kOk;
  long blocks = (n + 255) / 256;
  if (blocks > 1 << 20) blocks = 1 << 20;
  hipLaunchKernelGGL(stream_copy, dim3((unsigned)blocks), dim3(256), 0,
                     static_cast<hipStream_t>(stream), static_cast<float4*>(d_dst),
                     static_cast<const float4*>(d_src), n);
  return hipGetLastError() == hipSuccess ? kOk : kHipError;
}

}  // extern "C"
