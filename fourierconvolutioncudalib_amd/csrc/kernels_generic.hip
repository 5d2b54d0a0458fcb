// kernels_generic.hip -- runtime-length pass kernels (any FFT length that
// fits the LDS tile).  Each pass reads its tile from HBM once, transforms it
// in LDS and writes it back once.
//
// Layouts (C-order, x fastest):
//   real volume      [nz][ny][nx] float
//   half spectrum S  [nz][ny][nxp] float2, bins 0..nx/2 valid, nxp = nx/2+1
//                    rounded up to 16 bins (128-byte aligned column tiles)
//   kernel rows Kx   [kz][ky][nxp] float2 (x-transformed kernel rows)
//   kernel planes Kxy[kz][ny][nxp] float2 (x- and y-transformed kernel)
#include "fc_internal.hpp"
#include "fft_generic.hpp"
#include "kernels.hpp"

namespace fcamd {

// ---------------------------------------------------------------------------
// x pass, forward: two real rows a, b are packed as z = a + i*b, one complex
// FFT of length nx, then A[k] = (Z[k] + conj Z[-k])/2, B[k] = (Z[k] - conj
// Z[-k])/(2i) for k = 0..nx/2.  With `shift` the rows come from the compact
// kernel and are placed circularly so tap x lands at (x - kx/2) mod nx
// (src/convolution3Dfft.cu:150-158), zeros elsewhere.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
gk_xfwd(XPassArgs a) {
  extern __shared__ float2 smem[];
  const int L = a.fft.L;
  const int P = a.pairs_per_block;
  const int LP = a.fft.buf() + 1;  // odd row stride spreads rows over LDS banks
  float2* b0 = smem;
  float2* b1 = smem + (size_t)LP * P;
  const long pair0 = (long)blockIdx.x * P;

  for (int w = threadIdx.x; w < L * P; w += blockDim.x) {
    const int pos = w % L, c = w / L;
    const long ra = 2 * (pair0 + c), rb = ra + 1;
    float va = 0.f, vb = 0.f;
    int x = pos;
    bool ok = true;
    if (a.kshift) {
      x = pos + a.kcenter;
      if (x >= L) x -= L;
      ok = x < a.klen;
    }
    if (ok && a.vp.on && !a.kshift) {
      // padded image rows: read through the boundary maps
      bool okx;
      const int xs = vpad_src(a.vp, 0, pos, &okx);
      long sr;
      if (okx && ra < a.nrows && vpad_row(a.vp, ra, &sr)) va = a.src[sr * a.vp.n[0] + xs];
      if (okx && rb < a.nrows && vpad_row(a.vp, rb, &sr)) vb = a.src[sr * a.vp.n[0] + xs];
    } else if (ok) {
      if (ra < a.nrows) va = a.src[ra * a.src_row + x];
      if (rb < a.nrows) vb = a.src[rb * a.src_row + x];
    }
    b0[pos + c * LP] = make_float2(va, vb);
  }
  __syncthreads();
  const TileGeom g{P, 1, LP};
  float2* z = gfft_any<-1, false>(b0, b1, a.fft, g);

  const int nb = a.nbins;
  for (int w = threadIdx.x; w < nb * P; w += blockDim.x) {
    const int kb = w % nb, c = w / nb;
    const long ra = 2 * (pair0 + c), rb = ra + 1;
    const int km = kb == 0 ? 0 : L - kb;
    const float2 Z = z[kb + c * LP];
    const float2 Zm = z[km + c * LP];
    if (ra < a.nrows)
      a.dst[ra * a.dst_row + kb] = make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
    if (rb < a.nrows)
      a.dst[rb * a.dst_row + kb] = make_float2(0.5f * (Z.y + Zm.y), 0.5f * (Zm.x - Z.x));
  }
}

// ---------------------------------------------------------------------------
// x pass, inverse (C2R): rows a, b of the half spectrum are extended
// Hermitian (imaginary parts of the DC and Nyquist bins dropped, as a C2R
// does), packed as Z = A + i*B, inverse FFT; real part -> row a, imaginary
// part -> row b.  Unnormalised.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
gk_xinv(XPassArgs a) {
  extern __shared__ float2 smem[];
  const int L = a.fft.L;
  const int P = a.pairs_per_block;
  const int LP = a.fft.buf() + 1;
  float2* b0 = smem;
  float2* b1 = smem + (size_t)LP * P;
  const long pair0 = (long)blockIdx.x * P;
  const int nb = a.nbins;

  for (int w = threadIdx.x; w < L * P; w += blockDim.x) {
    const int pos = w % L, c = w / L;
    const long ra = 2 * (pair0 + c), rb = ra + 1;
    const bool mirror = pos >= nb;
    const int kb = mirror ? L - pos : pos;
    float2 A = make_float2(0.f, 0.f), B = make_float2(0.f, 0.f);
    if (ra < a.nrows) A = a.csrc[ra * a.src_row + kb];
    if (rb < a.nrows) B = a.csrc[rb * a.src_row + kb];
    if (kb == 0 || 2 * kb == L) {
      A.y = 0.f;
      B.y = 0.f;
    }
    if (mirror) {
      A.y = -A.y;
      B.y = -B.y;
    }
    b0[pos + c * LP] = make_float2(A.x - B.y, A.y + B.x);
  }
  __syncthreads();
  const TileGeom g{P, 1, LP};
  float2* z = gfft_any<+1, false>(b0, b1, a.fft, g);

  for (int w = threadIdx.x; w < L * P; w += blockDim.x) {
    const int pos = w % L, c = w / L;
    const long ra = 2 * (pair0 + c), rb = ra + 1;
    const float2 v = z[pos + c * LP];
    if (a.vp.on) {
      // cropped output: interior rows and columns only
      const int xo = pos - a.vp.lo[0];
      if (xo < 0 || xo >= a.vp.m[0]) continue;
      long orow;
      if (ra < a.nrows && vpad_out_row(a.vp, ra, &orow)) a.rdst[orow * a.vp.m[0] + xo] = v.x;
      if (rb < a.nrows && vpad_out_row(a.vp, rb, &orow)) a.rdst[orow * a.vp.m[0] + xo] = v.y;
      continue;
    }
    if (ra < a.nrows) a.rdst[ra * a.dst_row + pos] = v.x;
    if (rb < a.nrows) a.rdst[rb * a.dst_row + pos] = v.y;
  }
}

// ---------------------------------------------------------------------------
// column pass (y or z axis): a tile of TX consecutive bins by the full column
// length.  Source positions may be a circularly shifted, zero-padded compact
// array (the kernel's y rows) or the same dense spectrum (in place).
// ---------------------------------------------------------------------------
template <int DIR>
__device__ void col_body(const ColArgs& a) {
  extern __shared__ float2 smem[];
  const int L = a.fft.L;
  const int TX = a.tx;
  float2* b0 = smem;
  float2* b1 = smem + (size_t)a.fft.buf() * TX;
  const int x0 = blockIdx.x * TX;
  const long o = blockIdx.y;

  for (int w = threadIdx.x; w < L * TX; w += blockDim.x) {
    const int c = w % TX, pos = w / TX;
    int j = pos + a.center;
    if (j >= L) j -= L;
    float2 v = make_float2(0.f, 0.f);
    if (x0 + c < a.nbins && j < a.nvalid) v = a.src[o * a.src_outer + (long)j * a.src_elem + x0 + c];
    b0[pos * TX + c] = v;
  }
  __syncthreads();
  const TileGeom g{TX, TX, 1};
  float2* r = gfft_any<DIR, true>(b0, b1, a.fft, g);
  for (int w = threadIdx.x; w < L * TX; w += blockDim.x) {
    const int c = w % TX, pos = w / TX;
    if (x0 + c < a.nbins) a.dst[o * a.dst_outer + (long)pos * a.dst_elem + x0 + c] = r[pos * TX + c];
  }
}

__global__ void __launch_bounds__(256) gk_col_fwd(ColArgs a) { col_body<-1>(a); }
__global__ void __launch_bounds__(256) gk_col_inv(ColArgs a) { col_body<+1>(a); }

// ---------------------------------------------------------------------------
// fused z pass: forward z FFT of the image column tile, forward z FFT of the
// operand column (compact shifted kernel planes or a dense spectrum), complex
// product scaled by 1/N (src/convolution3Dfft.cu:41-62,531; conjugated
// operand for the adjoint), inverse z FFT, store in place.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
gk_zfused(ZArgs a) {
  extern __shared__ float2 smem[];
  const int L = a.fft.L;
  const int TX = a.tx;
  const size_t tile = (size_t)a.fft.buf() * TX;
  float2* b0 = smem;
  float2* b1 = smem + tile;
  float2* b2 = smem + 2 * tile;
  const int x0 = blockIdx.x * TX;
  const long y = blockIdx.y;
  const TileGeom g{TX, TX, 1};

  for (int w = threadIdx.x; w < L * TX; w += blockDim.x) {
    const int c = w % TX, pos = w / TX;
    float2 v = make_float2(0.f, 0.f);
    if (x0 + c < a.nbins) v = a.s[y * a.s_row + (long)pos * a.s_plane + x0 + c];
    b0[pos * TX + c] = v;
  }
  __syncthreads();
  float2* A = gfft_any<-1, true>(b0, b1, a.fft, g);
  float2* f0 = (A == b0) ? b1 : b0;  // the free image buffer

  for (int w = threadIdx.x; w < L * TX; w += blockDim.x) {
    const int c = w % TX, pos = w / TX;
    int j = pos + a.center;
    if (j >= L) j -= L;
    float2 v = make_float2(0.f, 0.f);
    if (x0 + c < a.nbins && j < a.nvalid) v = a.k[y * a.k_row + (long)j * a.k_plane + x0 + c];
    f0[pos * TX + c] = v;
  }
  __syncthreads();
  float2* K = gfft_any<-1, true>(f0, b2, a.fft, g);
  float2* f1 = (K == f0) ? b2 : f0;

  for (int w = threadIdx.x; w < L * TX; w += blockDim.x) {
    const float2 x = A[w], kk = K[w];
    const float2 p = a.conj ? cmulc(x, kk) : cmul(x, kk);
    A[w] = cscale(p, a.scale);
  }
  __syncthreads();
  float2* R = gfft_any<+1, true>(A, f1, a.fft, g);
  for (int w = threadIdx.x; w < L * TX; w += blockDim.x) {
    const int c = w % TX, pos = w / TX;
    int q = pos;
    if (a.out_planes > 0) {  // rotated subset (see ZArgs)
      q = pos + a.out_shift;
      if (q >= L) q -= L;
      if (q >= a.out_planes) continue;
    }
    if (x0 + c < a.nbins) a.s[y * a.s_row + (long)q * a.s_plane + x0 + c] = R[pos * TX + c];
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_gk_xfwd(const XPassArgs& a, long npairs, hipStream_t st) {
  const size_t lds = 2 * sizeof(float2) * (size_t)(a.fft.buf() + 1) * a.pairs_per_block;
  const unsigned grid = (unsigned)((npairs + a.pairs_per_block - 1) / a.pairs_per_block);
  hipLaunchKernelGGL(gk_xfwd, dim3(grid), dim3(256), lds, st, a);
}

void launch_gk_xinv(const XPassArgs& a, long npairs, hipStream_t st) {
  const size_t lds = 2 * sizeof(float2) * (size_t)(a.fft.buf() + 1) * a.pairs_per_block;
  const unsigned grid = (unsigned)((npairs + a.pairs_per_block - 1) / a.pairs_per_block);
  hipLaunchKernelGGL(gk_xinv, dim3(grid), dim3(256), lds, st, a);
}

void launch_gk_col(const ColArgs& a, int dir, long nouter, hipStream_t st) {
  const size_t lds = 2 * sizeof(float2) * (size_t)a.fft.buf() * a.tx;
  const dim3 grid((a.nbins + a.tx - 1) / a.tx, (unsigned)nouter);
  if (dir < 0)
    hipLaunchKernelGGL(gk_col_fwd, grid, dim3(256), lds, st, a);
  else
    hipLaunchKernelGGL(gk_col_inv, grid, dim3(256), lds, st, a);
}

void launch_gk_zfused(const ZArgs& a, long ny, hipStream_t st) {
  const size_t lds = 3 * sizeof(float2) * (size_t)a.fft.buf() * a.tx;
  const dim3 grid((a.nbins + a.tx - 1) / a.tx, (unsigned)ny);
  hipLaunchKernelGGL(gk_zfused, grid, dim3(256), lds, st, a);
}

void generic_kernels_allow_large_lds() {
  const int maxb = 160 * 1024;
  (void)hipFuncSetAttribute((const void*)gk_xfwd, hipFuncAttributeMaxDynamicSharedMemorySize, maxb);
  (void)hipFuncSetAttribute((const void*)gk_xinv, hipFuncAttributeMaxDynamicSharedMemorySize, maxb);
  (void)hipFuncSetAttribute((const void*)gk_col_fwd, hipFuncAttributeMaxDynamicSharedMemorySize, maxb);
  (void)hipFuncSetAttribute((const void*)gk_col_inv, hipFuncAttributeMaxDynamicSharedMemorySize, maxb);
  (void)hipFuncSetAttribute((const void*)gk_zfused, hipFuncAttributeMaxDynamicSharedMemorySize, maxb);
}

}  // namespace fcamd
