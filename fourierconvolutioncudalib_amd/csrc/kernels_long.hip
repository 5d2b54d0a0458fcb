// kernels_long.hip -- kernels of the long-axis path: transform lengths that
// exceed one LDS tile (x > 10239, y > 10240, z > 6826).  Such an axis of
// length L = L1 * L2 is transformed in four steps through HBM:
//   A  for every n2 < L2: an L1-point FFT over n1 of element n1*L2 + n2,
//      then the twiddle exp(-2 pi i k1 n2 / L) on output k1;
//   B  for every k1 < L1: an L2-point FFT over n2 of element k1*L2 + n2.
// Position k1*L2 + k2 then holds frequency k1 + L1*k2 (a transposed order,
// the same for the image and the kernel, so the pointwise product does not
// see it); the inverse runs B then A backwards and restores natural order.
// Every launch is a 1-D grid; each FFT runs in LDS on the runtime-length
// engine (fft_generic.hpp).
#include "fc_internal.hpp"
#include "fft_generic.hpp"
#include "kernels_long.hpp"

namespace fcamd {

__device__ __forceinline__ long lcol_base(const LColArgs& a, long q, bool src) {
  const long hi = src ? a.src_hi : a.dst_hi, lo = src ? a.src_lo : a.dst_lo;
  return a.outer_div > 0 ? (q / a.outer_div) * hi + (q % a.outer_div) * lo : q * hi;
}

// Column FFTs over a [outer][L][cols] view: a tile of `tx` consecutive
// columns by the whole length.  twiddle: 0 none; 1 forward step A (after the
// FFT, output k times exp(-2 pi i k n2 / tw_len)); 2 inverse step A (before
// the FFT, input k times exp(+2 pi i k n2 / tw_len)); n2 is the column index
// (tw_by_col) or the outer index modulo outer_div.
template <int DIR>
__device__ void lcol_body(const LColArgs& a) {
  extern __shared__ float2 smem[];
  const int L = a.fft.L, TX = a.tx;
  float2* b0 = smem;
  float2* b1 = smem + (size_t)a.fft.buf() * TX;
  const long ntx = (a.ncols + TX - 1) / TX;
  const long q = blockIdx.x / ntx;
  const long x0 = (blockIdx.x % ntx) * TX;
  const long sb = lcol_base(a, q, true), db = lcol_base(a, q, false);
  const long n2_outer = a.outer_div > 0 ? q % a.outer_div : 0;

  for (int w = threadIdx.x; w < L * TX; w += blockDim.x) {
    const int c = w % TX, pos = w / TX;
    float2 v = make_float2(0.f, 0.f);
    if (x0 + c < a.ncols) {
      v = a.src[sb + (long)pos * a.src_elem + (x0 + c) * a.src_col];
      if (a.twiddle == 2) {
        const long n2 = a.tw_by_col ? x0 + c : n2_outer;
        v = cmulc(v, a.tw_long[((long)pos * n2) % a.tw_len]);
      }
    }
    b0[pos * TX + c] = v;
  }
  __syncthreads();
  const TileGeom g{TX, TX, 1};
  float2* r = gfft_any<DIR, true>(b0, b1, a.fft, g);
  for (int w = threadIdx.x; w < L * TX; w += blockDim.x) {
    const int c = w % TX, pos = w / TX;
    if (x0 + c >= a.ncols) continue;
    float2 v = r[pos * TX + c];
    if (a.twiddle == 1) {
      const long n2 = a.tw_by_col ? x0 + c : n2_outer;
      v = cmul(v, a.tw_long[((long)pos * n2) % a.tw_len]);
    }
    a.dst[db + (long)pos * a.dst_elem + (x0 + c) * a.dst_col] = v;
  }
}

__global__ void __launch_bounds__(256) lk_col_fwd(LColArgs a) { lcol_body<-1>(a); }
__global__ void __launch_bounds__(256) lk_col_inv(LColArgs a) { lcol_body<+1>(a); }

// Two real rows per complex row: z[p][x] = v[2p][x] + i v[2p+1][x] (the
// missing partner of an odd last row reads zero).
__global__ void __launch_bounds__(256)
lk_xpack(const float* __restrict__ v, float2* __restrict__ z, long nrows, int nx) {
  const long npairs = (nrows + 1) / 2, n = npairs * nx;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const long p = i / nx;
    const int x = (int)(i % nx);
    const long ra = 2 * p, rb = ra + 1;
    z[i] = make_float2(v[ra * nx + x], rb < nrows ? v[rb * nx + x] : 0.f);
  }
}

// position of frequency f in the four-step order (identity when L1 == 0)
__device__ __forceinline__ long fpos(long f, int L1, int L2) {
  return L1 > 0 ? (f % L1) * L2 + f / L1 : f;
}

// Separation of the packed rows' spectra into the half-spectrum rows of both
// real rows: A[k] = (Z[k] + conj Z[-k]) / 2, B[k] = (Z[k] - conj Z[-k]) / 2i.
__global__ void __launch_bounds__(256)
lk_xsep(const float2* __restrict__ z, float2* __restrict__ S, long nrows, int nx, int nb,
        long srow, int L1, int L2) {
  const long npairs = (nrows + 1) / 2, n = npairs * nb;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const long p = i / nb;
    const int k = (int)(i % nb);
    const float2* zp = z + p * nx;
    const float2 Z = zp[fpos(k, L1, L2)];
    const float2 Zm = zp[fpos(k == 0 ? 0 : nx - k, L1, L2)];
    const long ra = 2 * p, rb = ra + 1;
    S[ra * srow + k] = make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
    if (rb < nrows) S[rb * srow + k] = make_float2(0.5f * (Z.y + Zm.y), 0.5f * (Zm.x - Z.x));
  }
}

// Hermitian extension of two half-spectrum rows packed as Z = A + iB over all
// nx frequencies (imaginary parts of DC and Nyquist dropped, as a C2R does),
// written in the four-step order for the inverse transform.
__global__ void __launch_bounds__(256)
lk_xunpack(const float2* __restrict__ S, float2* __restrict__ z, long nrows, int nx, int nb,
           long srow, int L1, int L2) {
  const long npairs = (nrows + 1) / 2, n = npairs * nx;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const long p = i / nx;
    const int f = (int)(i % nx);
    const bool mirror = f >= nb;
    const int k = mirror ? nx - f : f;
    const long ra = 2 * p, rb = ra + 1;
    float2 A = S[ra * srow + k];
    float2 B = rb < nrows ? S[rb * srow + k] : make_float2(0.f, 0.f);
    if (k == 0 || 2 * k == nx) {
      A.y = 0.f;
      B.y = 0.f;
    }
    if (mirror) {
      A.y = -A.y;
      B.y = -B.y;
    }
    z[p * nx + fpos(f, L1, L2)] = make_float2(A.x - B.y, A.y + B.x);
  }
}

// real part -> row 2p, imaginary part -> row 2p + 1
__global__ void __launch_bounds__(256)
lk_xreal(const float2* __restrict__ z, float* __restrict__ v, long nrows, int nx) {
  const long npairs = (nrows + 1) / 2, n = npairs * nx;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const long p = i / nx;
    const int x = (int)(i % nx);
    const float2 c = z[i];
    v[2 * p * nx + x] = c.x;
    if (2 * p + 1 < nrows) v[(2 * p + 1) * nx + x] = c.y;
  }
}

__global__ void __launch_bounds__(256) lk_fill_zero(float* __restrict__ v, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    v[i] = 0.f;
}

// the compact kernel placed circularly on the image grid: tap (z, y, x) at
// ((z - cz) mod nz, (y - cy) mod ny, (x - cx) mod nx)
__global__ void __launch_bounds__(256)
lk_scatter_kernel(const float* __restrict__ k, float* __restrict__ v, int nx, int ny, int nz,
                  int kx, int ky, int kz, int cx, int cy, int cz) {
  const long K = (long)kx * ky * kz;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < K;
       i += (long)gridDim.x * blockDim.x) {
    const int x = (int)(i % kx), y = (int)((i / kx) % ky), z = (int)(i / ((long)kx * ky));
    int X = x - cx, Y = y - cy, Z = z - cz;
    if (X < 0) X += nx;
    if (Y < 0) Y += ny;
    if (Z < 0) Z += nz;
    v[((long)Z * ny + Y) * nx + X] = k[i];
  }
}

// Bluestein's chirp-z for an axis of length L (no split into two tiles),
// over lines x(o, p, c) at o * hi + p * elem + c (o < nouter, p < L,
// c < ncols) and a scratch W[nouter][M][ncols]:
//   W(o, m, c) = x(o, m, c) * chirp[m] (m < L), 0 (L <= m < M)
//   W <- FFT_M(W) * bfft (four-step order), then the inverse FFT_M
//   x(o, k, c) = W(o, k, c) * chirp[k]   (k < L)
// with conj(chirp) and conj(bfft) for the inverse direction (dir > 0).
__device__ __forceinline__ float2 chirp_mul(float2 v, float2 w, int dir) {
  return dir < 0 ? cmul(v, w) : cmulc(v, w);
}

__global__ void __launch_bounds__(256)
lk_chirp_pre(const float2* __restrict__ x, long hi, long elem, long ncols, long nouter, int L,
             float2* __restrict__ W, int M, const float2* __restrict__ chirp, int dir) {
  const long n = nouter * M * ncols;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const long c = i % ncols, r = i / ncols;
    const int m = (int)(r % M);
    const long o = r / M;
    W[i] = m < L ? chirp_mul(x[o * hi + m * elem + c], chirp[m], dir) : make_float2(0.f, 0.f);
  }
}

__global__ void __launch_bounds__(256)
lk_chirp_mul(float2* __restrict__ W, long ncols, long n, int M, const float2* __restrict__ bfft,
             int dir) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    W[i] = chirp_mul(W[i], bfft[(i / ncols) % M], dir);
}

__global__ void __launch_bounds__(256)
lk_chirp_post(const float2* __restrict__ W, float2* __restrict__ x, long hi, long elem, long ncols,
              long nouter, int L, int M, const float2* __restrict__ chirp, int dir) {
  const long n = nouter * L * ncols;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const long c = i % ncols, r = i / ncols;
    const int k = (int)(r % L);
    const long o = r / L;
    x[o * hi + k * elem + c] = chirp_mul(W[(o * M + k) * ncols + c], chirp[k], dir);
  }
}

namespace {
unsigned ew_grid(long n) {
  long b = (n + 255) / 256;
  return (unsigned)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}
}  // namespace

void launch_lcol(const LColArgs& a, int dir, hipStream_t st) {
  const long ntx = (a.ncols + a.tx - 1) / a.tx;
  const long blocks = ntx * a.nouter;
  FC_REQUIRE(blocks >= 1 && blocks <= 0x7fffffffL, kUnsupportedSize, "too many column tiles");
  const size_t lds = 2 * sizeof(float2) * (size_t)a.fft.buf() * a.tx;
  if (dir < 0)
    hipLaunchKernelGGL(lk_col_fwd, dim3((unsigned)blocks), dim3(256), lds, st, a);
  else
    hipLaunchKernelGGL(lk_col_inv, dim3((unsigned)blocks), dim3(256), lds, st, a);
}

void launch_xpack(const float* v, float2* z, long nrows, int nx, hipStream_t st) {
  hipLaunchKernelGGL(lk_xpack, dim3(ew_grid((nrows + 1) / 2 * nx)), dim3(256), 0, st, v, z, nrows,
                     nx);
}
void launch_xsep(const float2* z, float2* S, long nrows, int nx, long srow, int L1, int L2,
                 hipStream_t st) {
  const int nb = nx / 2 + 1;
  hipLaunchKernelGGL(lk_xsep, dim3(ew_grid((nrows + 1) / 2 * nb)), dim3(256), 0, st, z, S, nrows,
                     nx, nb, srow, L1, L2);
}
void launch_xunpack(const float2* S, float2* z, long nrows, int nx, long srow, int L1, int L2,
                    hipStream_t st) {
  const int nb = nx / 2 + 1;
  hipLaunchKernelGGL(lk_xunpack, dim3(ew_grid((nrows + 1) / 2 * nx)), dim3(256), 0, st, S, z,
                     nrows, nx, nb, srow, L1, L2);
}
void launch_xreal(const float2* z, float* v, long nrows, int nx, hipStream_t st) {
  hipLaunchKernelGGL(lk_xreal, dim3(ew_grid((nrows + 1) / 2 * nx)), dim3(256), 0, st, z, v, nrows,
                     nx);
}
void launch_chirp_pre(const float2* x, long hi, long elem, long ncols, long nouter, int L,
                      float2* W, int M, const float2* chirp, int dir, hipStream_t st) {
  hipLaunchKernelGGL(lk_chirp_pre, dim3(ew_grid(nouter * M * ncols)), dim3(256), 0, st, x, hi, elem,
                     ncols, nouter, L, W, M, chirp, dir);
}
void launch_chirp_mul(float2* W, long ncols, long nouter, int M, const float2* bfft, int dir,
                      hipStream_t st) {
  const long n = nouter * M * ncols;
  hipLaunchKernelGGL(lk_chirp_mul, dim3(ew_grid(n)), dim3(256), 0, st, W, ncols, n, M, bfft, dir);
}
void launch_chirp_post(const float2* W, float2* x, long hi, long elem, long ncols, long nouter,
                       int L, int M, const float2* chirp, int dir, hipStream_t st) {
  hipLaunchKernelGGL(lk_chirp_post, dim3(ew_grid(nouter * L * ncols)), dim3(256), 0, st, W, x, hi,
                     elem, ncols, nouter, L, M, chirp, dir);
}
void launch_fill_zero(float* v, long n, hipStream_t st) {
  hipLaunchKernelGGL(lk_fill_zero, dim3(ew_grid(n)), dim3(256), 0, st, v, n);
}
void launch_scatter_kernel(const float* k, float* v, int nx, int ny, int nz, int kx, int ky, int kz,
                           int cx, int cy, int cz, hipStream_t st) {
  hipLaunchKernelGGL(lk_scatter_kernel, dim3(ew_grid((long)kx * ky * kz)), dim3(256), 0, st, k, v,
                     nx, ny, nz, kx, ky, kz, cx, cy, cz);
}

void long_kernels_allow_large_lds() {
  const int maxb = 160 * 1024;
  (void)hipFuncSetAttribute((const void*)lk_col_fwd, hipFuncAttributeMaxDynamicSharedMemorySize,
                            maxb);
  (void)hipFuncSetAttribute((const void*)lk_col_inv, hipFuncAttributeMaxDynamicSharedMemorySize,
                            maxb);
}

}  // namespace fcamd
