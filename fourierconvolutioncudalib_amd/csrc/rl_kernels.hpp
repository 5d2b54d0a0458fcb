// rl_kernels.hpp -- compile-time-length kernels of the fused multiview
// Richardson-Lucy iteration
//   psi <- psi * (1/V) sum_v corr(phi_v / max(conv(psi, P_v), eps), P_v)
// built on fft_static.hpp and the pass bodies of static_kernels_impl.hpp.
//
//   sk_zfan      one z tile of FFTxy(psi): FFTz once, then per view the
//                product with FFTz(P_v) and IFFTz into that view's spectrum
//   sk_xratio    per view, rows of the blurred estimate: x-C2R, ratio
//                phi / max(blur, eps) in registers, x-R2C back in place
//                (the ratio volume never reaches HBM)
//   sk_zsum      one z tile: sum over views of FFTz(ratio_v) * conj(FFTz(P_v)),
//                IFFTz (correlations add in the spectral domain)
//   sk_xupdate   x-C2R of the summed correlation with the epilogue
//                psi <- psi * acc / V
// (the x kernels also in flat forms, staging the workgroup's rows through
// LDS with coalesced HBM access; used when the real rows are dense)
#pragma once

#include <type_traits>

#include "static_kernels_impl.hpp"

namespace fcamd {

// minimum waves per SIMD asked of the two register-heaviest kernels (0: no
// request; the compiler then sizes registers for the code alone).  4 waves
// (the most their LDS allows) took the 10-iteration, 4-view 384 x 384 x 192
// run from 15.34 to 15.04 ms, although the ratio kernel then spills 8
// registers.
#ifndef FC_RL_XRATIO_WAVES
#define FC_RL_XRATIO_WAVES 4
#endif
#ifndef FC_RL_ZSUM_WAVES
#define FC_RL_ZSUM_WAVES 4
#endif
#define FC_RL_BOUNDS_0(n) __launch_bounds__(n)
#define FC_RL_BOUNDS_3(n) __launch_bounds__(n, 3)
#define FC_RL_BOUNDS_4(n) __launch_bounds__(n, 4)
#define FC_RL_BOUNDS_5(n) __launch_bounds__(n, 5)
#define FC_RL_BOUNDS_CAT(n, w) FC_RL_BOUNDS_##w(n)
#define FC_RL_BOUNDS(n, w) FC_RL_BOUNDS_CAT(n, w)

// loads the compact kernel column of view v (planes placed at (z - cz) mod
// L); positions are t + (compile-time constant) after unrolling, so each
// access is a per-thread base plus a uniform offset
template <int L, int MAXE>
__device__ __forceinline__ void load_kernel_column(float2 (&w)[MAXE], int t, bool ok,
                                                   const float2* kcol, long kplane, int kz, int cz) {
  const float2* kt = kcol + (long)(t + cz) * kplane;
  const float2* kw = kt - (long)L * kplane;
  sfft_for_inputs<L, false>(t, [&](int slot, int pos) {
    const bool wrap = pos + cz >= L;
    const int j = wrap ? pos + cz - L : pos + cz;
    float2 val = make_float2(0.f, 0.f);
    if (ok && j < kz) val = (wrap ? kw : kt)[(long)(pos - t) * kplane];
    w[slot] = val;
  });
}

// phi / d for d = max(blur, eps) > 0.  With eps a normal number (the
// launch-uniform case FAST) every d is one, and the hardware reciprocal (at
// most 1 ulp) times phi replaces the IEEE division sequence: one rounding
// more, some ten instructions fewer per sample.  A subnormal eps keeps the
// division.
constexpr float kMinNormal = 1.17549435e-38f;
template <bool FAST>
__device__ __forceinline__ float rl_div(float phi, float d) {
  if constexpr (FAST) return phi * __builtin_amdgcn_rcpf(d);
  return phi / d;
}

// column access at position pos through a base at position t
template <class T>
__device__ __forceinline__ T& col_at(T* base_t, int pos, int t, long stride) {
  return base_t[(long)(pos - t) * stride];
}

template <int L, int TX>
__global__ void __launch_bounds__(TX * threads_per_transform<L>())
sk_zfan(RLZArgs a) {
  using P = PlanOf<L, false>;
  __shared__ float2 lds[L * TX];
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  long y;
  int x;
  bool ok;
  column_tile_decode(a.nbins, TX, a.ny, blockIdx.x, c, &y, &x, &ok);
  const long base = ok ? y * a.s_row + x : 0;
  auto addr = [&](int pos) { return pos * TX + c; };
  float2 X[P::MAXE()];
  const float2* src_t = a.src + base + (long)t * a.s_plane;
  sfft_for_inputs<L, false>(t, [&](int slot, int pos) {
    float2 val = make_float2(0.f, 0.f);
    if (ok) val = col_at(src_t, pos, t, a.s_plane);
    X[slot] = val;
  });
  sfft_run<L, false, -1>(X, t, lds, addr, a.fft.stw);
  for (int v = 0; v < a.nviews; ++v) {
    float2 W[P::MAXE()];
    load_kernel_column<L>(W, t, ok, a.kplanes[v] + base, a.s_plane, a.kz[v], a.cz[v]);
    sfft_run<L, false, -1>(W, t, lds, addr, a.fft.stw);
    sfft_for_outputs<L, false>(t, [&](int slot, int) { W[slot] = cscale(cmul(X[slot], W[slot]), a.scale); });
    sfft_run<L, true, +1>(W, t, lds, addr, a.fft.stw);
    float2* out_t = a.view[v] + base + (long)t * a.s_plane;
    if (ok) sfft_for_outputs<L, true>(t, [&](int slot, int pos) { col_at(out_t, pos, t, a.s_plane) = W[slot]; });
  }
}

template <int L, int TX>
__global__ void FC_RL_BOUNDS(TX * threads_per_transform<L>(), FC_RL_ZSUM_WAVES)
sk_zsum(RLZArgs a) {
  using P = PlanOf<L, false>;
  __shared__ float2 lds[L * TX];
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  long y;
  int x;
  bool ok;
  column_tile_decode(a.nbins, TX, a.ny, blockIdx.x, c, &y, &x, &ok);
  const long base = ok ? y * a.s_row + x : 0;
  auto addr = [&](int pos) { return pos * TX + c; };
  float2 acc[P::MAXE()];
#pragma unroll
  for (int i = 0; i < P::MAXE(); ++i) acc[i] = make_float2(0.f, 0.f);
  for (int v = 0; v < a.nviews; ++v) {
    float2 X[P::MAXE()], W[P::MAXE()];
    const float2* in_t = a.view[v] + base + (long)t * a.s_plane;
    sfft_for_inputs<L, false>(t, [&](int slot, int pos) {
      float2 val = make_float2(0.f, 0.f);
      if (ok) val = col_at(in_t, pos, t, a.s_plane);
      X[slot] = val;
    });
    load_kernel_column<L>(W, t, ok, a.kplanes[v] + base, a.s_plane, a.kz[v], a.cz[v]);
    sfft_run<L, false, -1>(X, t, lds, addr, a.fft.stw);
    sfft_run<L, false, -1>(W, t, lds, addr, a.fft.stw);
    sfft_for_outputs<L, false>(t, [&](int slot, int) { acc[slot] = cadd(acc[slot], cmulc(X[slot], W[slot])); });
  }
  sfft_for_outputs<L, false>(t, [&](int slot, int) { acc[slot] = cscale(acc[slot], a.scale); });
  sfft_run<L, true, +1>(acc, t, lds, addr, a.fft.stw);
  float2* out_t = a.dst + base + (long)t * a.s_plane;
  if (ok) sfft_for_outputs<L, true>(t, [&](int slot, int pos) { col_at(out_t, pos, t, a.s_plane) = acc[slot]; });
}

// x-C2R (reversed plan: its outputs sit exactly at the forward plan's
// stage-0 input positions), ratio, x-R2C (forward plan), separation.
template <int L, int PAIRS>
__global__ void __launch_bounds__(PAIRS * threads_per_transform<L>())
sk_xratio(RLXArgs a) {
  using P = PlanOf<L, false>;
  constexpr int T = P::T, LP = XRow<L>::LP;
  __shared__ float2 lds[PAIRS * LP];
  const long groups = ((a.nrows + 1) / 2 + PAIRS - 1) / PAIRS;
  const int v = (int)(blockIdx.x / groups);
  const long blk = blockIdx.x - (long)v * groups;
  const int t = threadIdx.x % T, c = threadIdx.x / T;
  const long ra = 2 * (blk * PAIRS + c), rb = ra + 1;
  const bool oka = ra < a.nrows, okb = rb < a.nrows;
  float2* pa = a.spec[v] + ra * a.srow;
  float2* pb = a.spec[v] + rb * a.srow;
  const int nb = a.nbins;
  float2 z[P::MAXE()];
  sfft_for_inputs<L, true>(t, [&](int slot, int pos) {
    const bool mirror = pos >= nb;
    const int kb = mirror ? L - pos : pos;
    float2 A = make_float2(0.f, 0.f), B = make_float2(0.f, 0.f);
    if (oka) A = pa[kb];
    if (okb) B = pb[kb];
    if (kb == 0 || 2 * kb == L) {
      A.y = 0.f;
      B.y = 0.f;
    }
    if (mirror) {
      A.y = -A.y;
      B.y = -B.y;
    }
    z[slot] = make_float2(A.x - B.y, A.y + B.x);
  });
  float2* row = lds + c * LP;
  auto addr = [&](int pos) { return XRow<L>::at(pos); };
  sfft_run<L, true, +1>(z, t, row, addr, a.fft.stw);
  // real rows of the blurred estimate -> ratio rows (same positions)
  const float* fa = a.phi[v] + ra * a.rrow;
  const float* fb = a.phi[v] + rb * a.rrow;
  const float eps = a.eps;
  auto ratio = [&](auto fast) {
    sfft_for_outputs<L, true>(t, [&](int slot, int pos) {
      float ya = 0.f, yb = 0.f;
      constexpr bool F = decltype(fast)::value;
      if (oka) ya = rl_div<F>(fa[pos], fmaxf(z[slot].x, eps));
      if (okb) yb = rl_div<F>(fb[pos], fmaxf(z[slot].y, eps));
      z[slot] = make_float2(ya, yb);
    });
  };
  if (eps >= kMinNormal)
    ratio(std::true_type{});
  else
    ratio(std::false_type{});
  sfft_run<L, false, -1>(z, t, row, addr, a.fft.stw);
  sfft_for_outputs<L, false>(t, [&](int slot, int pos) { row[addr(pos)] = z[slot]; });
  __syncthreads();
  for (int kb = t; kb < nb; kb += T) {
    const float2 Z = row[addr(kb)];
    const float2 Zm = row[addr(kb == 0 ? 0 : L - kb)];
    if (oka) pa[kb] = make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
    if (okb) pb[kb] = make_float2(0.5f * (Z.y + Zm.y), 0.5f * (Zm.x - Z.x));
  }
}

template <int L, int PAIRS>
__global__ void __launch_bounds__(PAIRS * threads_per_transform<L>())
sk_xupdate(RLXArgs a) {
  using P = PlanOf<L, false>;
  constexpr int T = P::T, LP = XRow<L>::LP;
  __shared__ float2 lds[PAIRS * LP];
  const int t = threadIdx.x % T, c = threadIdx.x / T;
  const long ra = 2 * ((long)blockIdx.x * PAIRS + c), rb = ra + 1;
  const bool oka = ra < a.nrows, okb = rb < a.nrows;
  const float2* pa = a.spec[0] + ra * a.srow;
  const float2* pb = a.spec[0] + rb * a.srow;
  const int nb = a.nbins;
  float2 z[P::MAXE()];
  sfft_for_inputs<L, false>(t, [&](int slot, int pos) {
    const bool mirror = pos >= nb;
    const int kb = mirror ? L - pos : pos;
    float2 A = make_float2(0.f, 0.f), B = make_float2(0.f, 0.f);
    if (oka) A = pa[kb];
    if (okb) B = pb[kb];
    if (kb == 0 || 2 * kb == L) {
      A.y = 0.f;
      B.y = 0.f;
    }
    if (mirror) {
      A.y = -A.y;
      B.y = -B.y;
    }
    z[slot] = make_float2(A.x - B.y, A.y + B.x);
  });
  float2* row = lds + c * LP;
  auto addr = [&](int pos) { return XRow<L>::at(pos); };
  sfft_run<L, false, +1>(z, t, row, addr, a.fft.stw);
  float* qa = a.psi + ra * a.rrow;
  float* qb = a.psi + rb * a.rrow;
  const float iv = a.inv_views;
  sfft_for_outputs<L, false>(t, [&](int slot, int pos) {
    if (oka) qa[pos] = qa[pos] * z[slot].x * iv;
    if (okb) qb[pos] = qb[pos] * z[slot].y * iv;
  });
}

// Flat forms of the two x kernels: the workgroup's spectrum rows and real
// rows (phi, psi) move between HBM and LDS staging as coalesced runs, as in
// the flat x passes (xflat_* in static_kernels_impl.hpp).  Real rows must be
// dense (rrow == L).
template <int L, int PAIRS>
__global__ void FC_RL_BOUNDS(PAIRS * threads_per_transform<L>(), FC_RL_XRATIO_WAVES)
sk_xratio_flat(RLXArgs a) {
  using P = PlanOf<L, false>;
  constexpr int T = P::T, LP = XRow<L>::LP, SL = XStage<L>::SL, NB = XStage<L>::NB;
  constexpr int NT = PAIRS * T, MAXR = 2 * PAIRS;
  __shared__ float2 lds[xflat_lds_floats<L>(PAIRS) / 2];
  const long groups = ((a.nrows + 1) / 2 + PAIRS - 1) / PAIRS;
  const int v = (int)(blockIdx.x / groups);
  const long r0 = 2 * (blockIdx.x - (long)v * groups) * PAIRS;
  const int nr = a.nrows - r0 < MAXR ? (int)(a.nrows - r0) : MAXR;
  float2* spec = a.spec[v];
  xflat_load_bins<L, NT, MAXR>(spec + r0 * a.srow, a.srow, lds, nr);
  __syncthreads();
  const int t = threadIdx.x % T, c = threadIdx.x / T;
  const bool oka = 2 * c < nr, okb = 2 * c + 1 < nr;
  const float2* sa = lds + 2 * c * NB;
  const float2* sb = sa + NB;
  float2 z[P::MAXE()];
  // C2R on the reversed plan: its outputs sit where the forward plan reads
  sfft_for_inputs<L, true>(t, [&](int slot, int pos) {
    const bool mirror = pos >= NB;
    const int kb = mirror ? L - pos : pos;
    float2 A = oka ? sa[kb] : make_float2(0.f, 0.f);
    float2 B = okb ? sb[kb] : make_float2(0.f, 0.f);
    if (kb == 0 || 2 * kb == L) {
      A.y = 0.f;
      B.y = 0.f;
    }
    if (mirror) {
      A.y = -A.y;
      B.y = -B.y;
    }
    z[slot] = make_float2(A.x - B.y, A.y + B.x);
  });
  __syncthreads();
  float2* row = lds + c * LP;
  auto addr = [&](int pos) { return XRow<L>::at(pos); };
  sfft_run<L, true, +1>(z, t, row, addr, a.fft.stw);
  // phi rows through LDS; ratio at the transform's output positions
  float* st = reinterpret_cast<float*>(lds);
  xflat_load_rows<L, NT, MAXR>(a.phi[v] + r0 * L, st, nr);
  __syncthreads();
  const float eps = a.eps;
  auto ratio = [&](auto fast) {
    constexpr bool F = decltype(fast)::value;
    sfft_for_outputs<L, true>(t, [&](int slot, int pos) {
      const float ya = oka ? rl_div<F>(st[2 * c * SL + pos], fmaxf(z[slot].x, eps)) : 0.f;
      const float yb = okb ? rl_div<F>(st[(2 * c + 1) * SL + pos], fmaxf(z[slot].y, eps)) : 0.f;
      z[slot] = make_float2(ya, yb);
    });
  };
  if (eps >= kMinNormal)
    ratio(std::true_type{});
  else
    ratio(std::false_type{});
  __syncthreads();
  sfft_run<L, false, -1>(z, t, row, addr, a.fft.stw);
  sfft_for_outputs<L, false>(t, [&](int slot, int pos) { row[addr(pos)] = z[slot]; });
  __syncthreads();
  // Hermitian split back into the view's spectrum rows, bin-contiguous
  constexpr int IT = (MAXR * NB + NT - 1) / NT;
  const int n = nr * NB;
#pragma unroll
  for (int j = 0; j < IT; ++j) {
    const int i = threadIdx.x + j * NT;
    if (i < n) {
      const int r = i / NB, kb = i - r * NB;
      const float2* pr = lds + (r >> 1) * LP;
      const float2 Z = pr[addr(kb)];
      const float2 Zm = pr[addr(kb == 0 ? 0 : L - kb)];
      spec[(r0 + r) * a.srow + kb] = (r & 1) ? make_float2(0.5f * (Z.y + Zm.y), 0.5f * (Zm.x - Z.x))
                                            : make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
    }
  }
}

template <int L, int PAIRS>
__global__ void __launch_bounds__(PAIRS * threads_per_transform<L>())
sk_xupdate_flat(RLXArgs a) {
  using P = PlanOf<L, false>;
  constexpr int T = P::T, LP = XRow<L>::LP, SL = XStage<L>::SL, NB = XStage<L>::NB;
  constexpr int NT = PAIRS * T, MAXR = 2 * PAIRS;
  __shared__ float2 lds[xflat_lds_floats<L>(PAIRS) / 2];
  const long r0 = 2 * (long)blockIdx.x * PAIRS;
  const int nr = a.nrows - r0 < MAXR ? (int)(a.nrows - r0) : MAXR;
  xflat_load_bins<L, NT, MAXR>(a.spec[0] + r0 * a.srow, a.srow, lds, nr);
  __syncthreads();
  const int t = threadIdx.x % T, c = threadIdx.x / T;
  const bool oka = 2 * c < nr, okb = 2 * c + 1 < nr;
  const float2* sa = lds + 2 * c * NB;
  const float2* sb = sa + NB;
  float2 z[P::MAXE()];
  sfft_for_inputs<L, false>(t, [&](int slot, int pos) {
    const bool mirror = pos >= NB;
    const int kb = mirror ? L - pos : pos;
    float2 A = oka ? sa[kb] : make_float2(0.f, 0.f);
    float2 B = okb ? sb[kb] : make_float2(0.f, 0.f);
    if (kb == 0 || 2 * kb == L) {
      A.y = 0.f;
      B.y = 0.f;
    }
    if (mirror) {
      A.y = -A.y;
      B.y = -B.y;
    }
    z[slot] = make_float2(A.x - B.y, A.y + B.x);
  });
  __syncthreads();
  float2* row = lds + c * LP;
  auto addr = [&](int pos) { return XRow<L>::at(pos); };
  sfft_run<L, false, +1>(z, t, row, addr, a.fft.stw);
  // psi rows through LDS: psi <- psi * acc / V at the output positions
  float* st = reinterpret_cast<float*>(lds);
  float* psi = a.psi + r0 * L;
  xflat_load_rows<L, NT, MAXR>(psi, st, nr);
  __syncthreads();
  const float iv = a.inv_views;
  sfft_for_outputs<L, false>(t, [&](int slot, int pos) {
    if (oka) st[2 * c * SL + pos] *= z[slot].x * iv;
    if (okb) st[(2 * c + 1) * SL + pos] *= z[slot].y * iv;
  });
  __syncthreads();
  xflat_store_rows<L, NT, MAXR>(psi, st, nr);
}

#ifndef FC_RL_XFLAT
#define FC_RL_XFLAT 1
#endif

template <int L>
void rl_launch_zfan(const RLZArgs& a, long ny, hipStream_t st) {
  constexpr int TX = fit_tile<L, 16>();
  RLZArgs b = a;
  b.ny = ny;
  hipLaunchKernelGGL((sk_zfan<L, TX>), dim3((unsigned)column_tiles(a.nbins, TX, ny)),
                     dim3(TX * threads_per_transform<L>()), 0, st, b);
}

template <int L>
void rl_launch_zsum(const RLZArgs& a, long ny, hipStream_t st) {
  constexpr int TX = fit_tile<L, 16>();
  RLZArgs b = a;
  b.ny = ny;
  hipLaunchKernelGGL((sk_zsum<L, TX>), dim3((unsigned)column_tiles(a.nbins, TX, ny)),
                     dim3(TX * threads_per_transform<L>()), 0, st, b);
}

// the flat forms need dense real rows; the per-pair forms take any stride
template <int L>
void rl_launch_xratio(const RLXArgs& a, int nviews, hipStream_t st) {
  if (FC_RL_XFLAT && a.rrow == L) {
    constexpr int P = flat_pairs<L, 256>();
    const long groups = ((a.nrows + 1) / 2 + P - 1) / P;
    hipLaunchKernelGGL((sk_xratio_flat<L, P>), dim3((unsigned)(groups * nviews)),
                       dim3(P * threads_per_transform<L>()), 0, st, a);
    return;
  }
  constexpr int P = fit_pairs<L, pairs_for<L, 256>()>();
  const long groups = ((a.nrows + 1) / 2 + P - 1) / P;
  hipLaunchKernelGGL((sk_xratio<L, P>), dim3((unsigned)(groups * nviews)),
                     dim3(P * threads_per_transform<L>()), 0, st, a);
}

template <int L>
void rl_launch_xupdate(const RLXArgs& a, hipStream_t st) {
  if (FC_RL_XFLAT && a.rrow == L) {
    constexpr int P = flat_pairs<L, 256>();
    const long groups = ((a.nrows + 1) / 2 + P - 1) / P;
    hipLaunchKernelGGL((sk_xupdate_flat<L, P>), dim3((unsigned)groups),
                       dim3(P * threads_per_transform<L>()), 0, st, a);
    return;
  }
  constexpr int P = fit_pairs<L, pairs_for<L, 256>()>();
  const long groups = ((a.nrows + 1) / 2 + P - 1) / P;
  hipLaunchKernelGGL((sk_xupdate<L, P>), dim3((unsigned)groups), dim3(P * threads_per_transform<L>()),
                     0, st, a);
}

}  // namespace fcamd
