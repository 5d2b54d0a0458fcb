// static_kernels_impl.hpp -- compile-time-length pass kernels built on
// fft_static.hpp.  Instantiated per length by static_inst.hip; the driver
// reaches them through the registry in static_kernels.hip.
#pragma once

#include "fft_static.hpp"
#include "kernels.hpp"
#include "static_registry.hpp"

namespace fcamd {

// bins per column tile: 16 (one 128-byte line of float2) when the
// workgroup stays within 1024 threads and 64 KiB of LDS, else 8, 4, ...
template <int L>
constexpr int col_tile_static() {
  constexpr int T = PlanOf<L, false>::T;
  int tx = 16;
  while (tx > 1 && (tx * T > 1024 || (long)tx * L * 8 > 65536)) tx /= 2;
  return tx;
}

// row pairs per workgroup of the x passes (about 256 threads)
template <int L>
constexpr int x_pairs_static() {
  constexpr int T = PlanOf<L, false>::T;
  int p = 256 / T;
  if (p < 1) p = 1;
  while (p > 1 && (long)p * (L + L / 8 + 1) * 8 > 65536) p /= 2;
  return p;
}

template <int L>
constexpr int x_threads_static() {
  return x_pairs_static<L>() * PlanOf<L, false>::T;
}
template <int L>
constexpr int col_threads_static() {
  return col_tile_static<L>() * PlanOf<L, false>::T;
}

// padded LDS row of the x passes: one float2 of padding per 8 keeps the
// stride-8 and stride-Ns stores of the early stages off a single bank set
template <int L>
struct XRow {
  static constexpr int LP = L + L / 8 + 1;
  __device__ static int at(int pos) { return pos + (pos >> 3); }
};

// ---------------------------------------------------------------------------
// x pass forward: two real rows per complex FFT (see kernels_generic.hip)
// ---------------------------------------------------------------------------
template <int L>
__global__ void __launch_bounds__(x_threads_static<L>())
sk_xfwd(XPassArgs a) {
  using P = PlanOf<L, false>;
  constexpr int T = P::T, PAIRS = x_pairs_static<L>(), LP = XRow<L>::LP;
  __shared__ float2 lds[PAIRS * LP];
  const int t = threadIdx.x % T, c = threadIdx.x / T;
  const long ra = 2 * ((long)blockIdx.x * PAIRS + c), rb = ra + 1;
  const bool oka = ra < a.nrows, okb = rb < a.nrows;
  const float* pa = a.src + ra * a.src_row;
  const float* pb = a.src + rb * a.src_row;
  float2 v[P::MAXE()];
  sfft_for_inputs<L, false>(t, [&](int slot, int pos) {
    float xa = 0.f, xb = 0.f;
    if (oka) xa = pa[pos];
    if (okb) xb = pb[pos];
    v[slot] = make_float2(xa, xb);
  });
  float2* row = lds + c * LP;
  auto addr = [&](int pos) { return XRow<L>::at(pos); };
  sfft_run<L, false, -1>(v, t, row, addr, a.fft.tw);
  sfft_for_outputs<L, false>(t, [&](int slot, int pos) { row[addr(pos)] = v[slot]; });
  __syncthreads();
  float2* da = a.dst + ra * a.dst_row;
  float2* db = a.dst + rb * a.dst_row;
  for (int kb = t; kb < a.nbins; kb += T) {
    const float2 Z = row[addr(kb)];
    const float2 Zm = row[addr(kb == 0 ? 0 : L - kb)];
    if (oka) da[kb] = make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
    if (okb) db[kb] = make_float2(0.5f * (Z.y + Zm.y), 0.5f * (Zm.x - Z.x));
  }
}

// ---------------------------------------------------------------------------
// x pass inverse (C2R), two rows per complex FFT
// ---------------------------------------------------------------------------
template <int L>
__global__ void __launch_bounds__(x_threads_static<L>())
sk_xinv(XPassArgs a) {
  using P = PlanOf<L, false>;
  constexpr int T = P::T, PAIRS = x_pairs_static<L>(), LP = XRow<L>::LP;
  __shared__ float2 lds[PAIRS * LP];
  const int t = threadIdx.x % T, c = threadIdx.x / T;
  const long ra = 2 * ((long)blockIdx.x * PAIRS + c), rb = ra + 1;
  const bool oka = ra < a.nrows, okb = rb < a.nrows;
  const float2* pa = a.csrc + ra * a.src_row;
  const float2* pb = a.csrc + rb * a.src_row;
  const int nb = a.nbins;
  float2 v[P::MAXE()];
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
    v[slot] = make_float2(A.x - B.y, A.y + B.x);
  });
  float2* row = lds + c * LP;
  auto addr = [&](int pos) { return XRow<L>::at(pos); };
  sfft_run<L, false, +1>(v, t, row, addr, a.fft.tw);
  float* oa = a.rdst + ra * a.dst_row;
  float* ob = a.rdst + rb * a.dst_row;
  sfft_for_outputs<L, false>(t, [&](int slot, int pos) {
    if (oka) oa[pos] = v[slot].x;
    if (okb) ob[pos] = v[slot].y;
  });
}

// ---------------------------------------------------------------------------
// column pass (y): tile of TX bins x L positions
// ---------------------------------------------------------------------------
template <int L, int DIR>
__global__ void __launch_bounds__(col_threads_static<L>())
sk_col(ColArgs a) {
  using P = PlanOf<L, false>;
  constexpr int TX = col_tile_static<L>();
  __shared__ float2 lds[L * TX];
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  const int x = blockIdx.x * TX + c;
  const bool ok = x < a.nbins;
  const long o = blockIdx.y;
  const float2* src = a.src + o * a.src_outer + x;
  float2 v[P::MAXE()];
  sfft_for_inputs<L, false>(t, [&](int slot, int pos) {
    int j = pos + a.center;
    if (j >= L) j -= L;
    float2 val = make_float2(0.f, 0.f);
    if (ok && j < a.nvalid) val = src[(long)j * a.src_elem];
    v[slot] = val;
  });
  auto addr = [&](int pos) { return pos * TX + c; };
  sfft_run<L, false, DIR>(v, t, lds, addr, a.fft.tw);
  float2* dst = a.dst + o * a.dst_outer + x;
  if (ok) sfft_for_outputs<L, false>(t, [&](int slot, int pos) { dst[(long)pos * a.dst_elem] = v[slot]; });
}

// ---------------------------------------------------------------------------
// fused z pass: forward image column, forward kernel column, product,
// inverse (reversed plan) -- all in registers between the FFTs
// ---------------------------------------------------------------------------
template <int L>
__global__ void __launch_bounds__(col_threads_static<L>())
sk_zfused(ZArgs a) {
  using P = PlanOf<L, false>;
  constexpr int TX = col_tile_static<L>();
  __shared__ float2 lds[L * TX];
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  const int x = blockIdx.x * TX + c;
  const bool ok = x < a.nbins;
  const long y = blockIdx.y;
  float2* col = a.s + y * a.s_row + x;
  const float2* kcol = a.k + y * a.k_row + x;
  auto addr = [&](int pos) { return pos * TX + c; };

  float2 v[P::MAXE()];
  sfft_for_inputs<L, false>(t, [&](int slot, int pos) {
    float2 val = make_float2(0.f, 0.f);
    if (ok) val = col[(long)pos * a.s_plane];
    v[slot] = val;
  });
  float2 w[P::MAXE()];
  sfft_for_inputs<L, false>(t, [&](int slot, int pos) {
    int j = pos + a.center;
    if (j >= L) j -= L;
    float2 val = make_float2(0.f, 0.f);
    if (ok && j < a.nvalid) val = kcol[(long)j * a.k_plane];
    w[slot] = val;
  });
  // every sfft_run ends with a barrier after its last LDS read, so the
  // next transform may reuse the tile
  sfft_run<L, false, -1>(v, t, lds, addr, a.fft.tw);
  sfft_run<L, false, -1>(w, t, lds, addr, a.fft.tw);
  const float sc = a.scale;
  const bool cj = a.conj != 0;
  sfft_for_outputs<L, false>(t, [&](int slot, int) {
    const float2 p = cj ? cmulc(v[slot], w[slot]) : cmul(v[slot], w[slot]);
    v[slot] = cscale(p, sc);
  });
  if constexpr (P::S > 1) __syncthreads();
  sfft_run<L, true, +1>(v, t, lds, addr, a.fft.tw);
  if (ok) sfft_for_outputs<L, true>(t, [&](int slot, int pos) { col[(long)pos * a.s_plane] = v[slot]; });
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int L>
void sk_launch_xfwd(const XPassArgs& a, long npairs, hipStream_t st) {
  constexpr int PAIRS = x_pairs_static<L>();
  const unsigned grid = (unsigned)((npairs + PAIRS - 1) / PAIRS);
  hipLaunchKernelGGL(sk_xfwd<L>, dim3(grid), dim3(PAIRS * PlanOf<L, false>::T), 0, st, a);
}

template <int L>
void sk_launch_xinv(const XPassArgs& a, long npairs, hipStream_t st) {
  constexpr int PAIRS = x_pairs_static<L>();
  const unsigned grid = (unsigned)((npairs + PAIRS - 1) / PAIRS);
  hipLaunchKernelGGL(sk_xinv<L>, dim3(grid), dim3(PAIRS * PlanOf<L, false>::T), 0, st, a);
}

template <int L>
void sk_launch_col(const ColArgs& a, int dir, long nouter, hipStream_t st) {
  constexpr int TX = col_tile_static<L>();
  const dim3 grid((a.nbins + TX - 1) / TX, (unsigned)nouter);
  const dim3 block(TX * PlanOf<L, false>::T);
  if (dir < 0)
    hipLaunchKernelGGL((sk_col<L, -1>), grid, block, 0, st, a);
  else
    hipLaunchKernelGGL((sk_col<L, +1>), grid, block, 0, st, a);
}

template <int L>
void sk_launch_zfused(const ZArgs& a, long ny, hipStream_t st) {
  constexpr int TX = col_tile_static<L>();
  const dim3 grid((a.nbins + TX - 1) / TX, (unsigned)ny);
  hipLaunchKernelGGL(sk_zfused<L>, grid, dim3(TX * PlanOf<L, false>::T), 0, st, a);
}

template <int L>
StaticEntry make_static_entry() {
  StaticEntry e;
  e.xfwd = &sk_launch_xfwd<L>;
  e.xinv = &sk_launch_xinv<L>;
  e.col = &sk_launch_col<L>;
  e.zfused = &sk_launch_zfused<L>;
  return e;
}

}  // namespace fcamd
