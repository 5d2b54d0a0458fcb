// fft_generic.hpp -- runtime-length Stockham FFT over a tile held in LDS.
//
// Handles any length whose plan fits the LDS tile: radices 2,3,4,5,7,8,16 run
// as butterflies, any other prime factor p runs as a direct p-point DFT stage
// (one output per work item, p complex MACs each).  Used for lengths that
// have no compile-time kernel and as the correctness baseline for those.
//
// Tile addressing: element (position p, transform c) lives at
// buf[p * ps + c * cs].  Work items are spread over threads with the
// transform index fastest (CFAST, column tiles) or the position fastest
// (row tiles) so that consecutive lanes touch consecutive LDS words.
#pragma once

#include "fc_internal.hpp"
#include "fft_radix.hpp"

namespace fcamd {

struct TileGeom {
  int C;   // transforms in the tile
  int ps;  // LDS stride between positions
  int cs;  // LDS stride between transforms
};

template <int R, int DIR, bool CFAST>
__device__ __forceinline__ void gstage_small(const float2* __restrict__ in, float2* __restrict__ out,
                                             int L, int Ns, const float2* __restrict__ tw,
                                             const TileGeom g) {
  const int M = L / R;
  const int twstep = L / (Ns * R);
  const int total = M * g.C;
  for (int w = threadIdx.x; w < total; w += blockDim.x) {
    int c, b;
    if (CFAST) { c = w % g.C; b = w / g.C; } else { b = w % M; c = w / M; }
    const int kk = b % Ns;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = in[(b + r * M) * g.ps + c * g.cs];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twiddle<DIR>(tw, r * kk * twstep));
    }
    Dft<R, DIR>::run(v);
    const int base = (b - kk) * R + kk;
#pragma unroll
    for (int r = 0; r < R; ++r) out[(base + r * Ns) * g.ps + c * g.cs] = v[r];
  }
}

template <int DIR, bool CFAST>
__device__ __forceinline__ void gstage_prime(const float2* __restrict__ in, float2* __restrict__ out,
                                             int R, int L, int Ns, const float2* __restrict__ tw,
                                             const TileGeom g) {
  const int M = L / R;
  const int twstep = L / (Ns * R);
  const int total = L * g.C;
  for (int w = threadIdx.x; w < total; w += blockDim.x) {
    int c, q;
    if (CFAST) { c = w % g.C; q = w / g.C; } else { q = w % L; c = w / L; }
    const int kk = q % Ns;
    const int m = (q / Ns) % R;
    const int b = (q / (Ns * R)) * Ns + kk;
    float2 acc = in[b * g.ps + c * g.cs];
    int rm = 0;  // (r*m) mod R
    for (int r = 1; r < R; ++r) {
      float2 x = in[(b + r * M) * g.ps + c * g.cs];
      x = cmul(x, twiddle<DIR>(tw, r * kk * twstep));
      rm += m;
      if (rm >= R) rm -= R;
      float2 wr = twiddle<DIR>(tw, rm * M);
      acc = cadd(acc, cmul(x, wr));
    }
    out[q * g.ps + c * g.cs] = acc;
  }
}

// Runs all stages; returns the buffer holding the result (a or b).
// Ends with a barrier, so the result may be read by any thread.
template <int DIR, bool CFAST>
__device__ float2* gfft(float2* a, float2* b, const FftDesc& d, const TileGeom g) {
  int Ns = 1;
  for (int s = 0; s < d.nst; ++s) {
    const int R = d.rad[s];
    switch (R) {
      case 2: gstage_small<2, DIR, CFAST>(a, b, d.L, Ns, d.tw, g); break;
      case 3: gstage_small<3, DIR, CFAST>(a, b, d.L, Ns, d.tw, g); break;
      case 4: gstage_small<4, DIR, CFAST>(a, b, d.L, Ns, d.tw, g); break;
      case 5: gstage_small<5, DIR, CFAST>(a, b, d.L, Ns, d.tw, g); break;
      case 7: gstage_small<7, DIR, CFAST>(a, b, d.L, Ns, d.tw, g); break;
      case 8: gstage_small<8, DIR, CFAST>(a, b, d.L, Ns, d.tw, g); break;
      case 16: gstage_small<16, DIR, CFAST>(a, b, d.L, Ns, d.tw, g); break;
      default: gstage_prime<DIR, CFAST>(a, b, R, d.L, Ns, d.tw, g); break;
    }
    __syncthreads();
    float2* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

// Whole transform of the tile whatever the plan: gfft, or for a Bluestein
// descriptor (d.M > 0) the chirp-z form
//   X[k] = c_k sum_n (x_n c_n) conj-chirp(k - n),  c = chirp (conj for DIR > 0)
// as a circular convolution of length M: premultiply and zero-pad, FFT_M,
// times bfft (conj for DIR > 0; the sequence is even), inverse FFT_M,
// postmultiply.  The tile buffers hold d.buf() positions; the caller loads
// positions [0, L) of `a`, and the result's positions [0, L) are valid.
template <int DIR, bool CFAST>
__device__ float2* gfft_any(float2* a, float2* b, const FftDesc& d, const TileGeom g) {
  if (d.M == 0) return gfft<DIR, CFAST>(a, b, d, g);
  const int L = d.L, M = d.M;
  for (int w = threadIdx.x; w < M * g.C; w += blockDim.x) {
    int c, p;
    if (CFAST) { c = w % g.C; p = w / g.C; } else { p = w % M; c = w / M; }
    float2& e = a[p * g.ps + c * g.cs];
    e = p < L ? (DIR < 0 ? cmul(e, d.chirp[p]) : cmulc(e, d.chirp[p])) : make_float2(0.f, 0.f);
  }
  __syncthreads();
  FftDesc dm = d;
  dm.L = M;
  float2* r = gfft<-1, CFAST>(a, b, dm, g);
  for (int w = threadIdx.x; w < M * g.C; w += blockDim.x) {
    int c, k;
    if (CFAST) { c = w % g.C; k = w / g.C; } else { k = w % M; c = w / M; }
    float2& e = r[k * g.ps + c * g.cs];
    e = DIR < 0 ? cmul(e, d.bfft[k]) : cmulc(e, d.bfft[k]);
  }
  __syncthreads();
  r = gfft<+1, CFAST>(r, r == a ? b : a, dm, g);
  for (int w = threadIdx.x; w < L * g.C; w += blockDim.x) {
    int c, k;
    if (CFAST) { c = w % g.C; k = w / g.C; } else { k = w % L; c = w / L; }
    float2& e = r[k * g.ps + c * g.cs];
    e = DIR < 0 ? cmul(e, d.chirp[k]) : cmulc(e, d.chirp[k]);
  }
  __syncthreads();
  return r;
}

}  // namespace fcamd
