// static_kernels_impl.hpp -- compile-time-length pass kernels built on
// fft_static.hpp, in several variants per pass.  Instantiated per length by
// static_inst.hip; the driver reaches them through the registry in
// static_kernels.hip.
//
// Variants:
//   x forward  "w<N>"            about N threads (N/T row pairs) per workgroup
//              "w<N>-flat"       rows staged through LDS with workgroup-wide
//                                coalesced loads and stores
//   x inverse  "w<N>", "w<N>-flat"
//              "w256-ldsmirror"  Hermitian extension formed in LDS
//   column     "t<TX>"           one TX-bin tile per workgroup
//              "t<TX>-pipe"      persistent workgroups; the next tile's loads
//                                are issued before the current tile's FFT
//   fused z    "t<TX>-fft"       kernel column by the same FFT as the image
//              "t<TX>-dft"       kernel spectrum by direct DFT over its kz
//                                taps straight into the last-stage registers
// and one batched form per pass (sb_*) whose workgroups take their
// (argument record, tile) from a device task table.
#pragma once

#include "fft_static.hpp"
#include "kernels.hpp"
#include "static_registry.hpp"

namespace fcamd {

template <int L>
constexpr int threads_per_transform() {
  return PlanOf<L, false>::T;
}

// largest tile <= want that keeps a workgroup within 1024 threads and
// 64 KiB of LDS
template <int L, int WANT>
constexpr int fit_tile() {
  int tx = WANT;
  while (tx > 1 && (tx * threads_per_transform<L>() > 1024 || (long)tx * L * 8 > 65536)) tx /= 2;
  return tx;
}

// row pairs per x-pass workgroup for about TGT threads (T threads per row
// pair), within the LDS and thread limits of fit_pairs
template <int L, int TGT>
constexpr int pairs_for() {
  const int t = threads_per_transform<L>();
  return t >= TGT ? 1 : TGT / t;
}

template <int L, int WANT>
constexpr int fit_pairs() {
  int p = WANT;
  while (p > 1 && (p * threads_per_transform<L>() > 1024 || (long)p * (L + L / 8 + 1) * 8 > 65536))
    p /= 2;
  return p;
}


// padded LDS row of the x passes: one float2 of padding per 8 keeps the
// stride-8 and stride-Ns stores of the early stages off a single bank set
template <int L>
struct XRow {
  static constexpr int LP = L + L / 8 + 1;
  __device__ static int at(int pos) { return pos + (pos >> 3); }
};

// ---------------------------------------------------------------------------
// x pass forward: two real rows per complex FFT (see kernels_generic.hip).
// The first k_blocks workgroups transform the compact kernel's rows, placed
// circularly so tap x lands at (x - kx/2) mod nx.
// ---------------------------------------------------------------------------
template <int L, int PAIRS>
__device__ __forceinline__ void xfwd_body(const XPassArgs& a, long blk, float2* lds) {
  using P = PlanOf<L, false>;
  constexpr int T = P::T, LP = XRow<L>::LP;
  const int t = threadIdx.x % T, c = threadIdx.x / T;
  const long ra = 2 * (blk * PAIRS + c), rb = ra + 1;
  const bool oka = ra < a.nrows, okb = rb < a.nrows;
  const float* pa = a.src + ra * a.src_row;
  const float* pb = a.src + rb * a.src_row;
  const int kc = a.kcenter, kl = a.klen;
  float2 v[P::MAXE()];
  sfft_for_inputs<L, false>(t, [&](int slot, int pos) {
    int x = pos + kc;
    if (x >= L) x -= L;
    float xa = 0.f, xb = 0.f;
    if (x < kl) {
      if (oka) xa = pa[x];
      if (okb) xb = pb[x];
    }
    v[slot] = make_float2(xa, xb);
  });
  float2* row = lds + c * LP;
  auto addr = [&](int pos) { return XRow<L>::at(pos); };
  sfft_run<L, false, -1, 9>(v, t, row, addr, a.fft.stw);
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

// Image rows of a virtually padded volume (VPad in kernels.hpp): row and
// element indices go through the boundary maps, zeros where undefined.
template <int L, int PAIRS>
__device__ __forceinline__ void xfwd_vpad_body(const XPassArgs& a, long blk, float2* lds) {
  using P = PlanOf<L, false>;
  constexpr int T = P::T, LP = XRow<L>::LP;
  const int t = threadIdx.x % T, c = threadIdx.x / T;
  const long ra = 2 * (blk * PAIRS + c), rb = ra + 1;
  const VPad& vp = a.vp;
  long sa = 0, sb = 0;
  const bool oka = ra < a.nrows && vpad_row(vp, ra, &sa);
  const bool okb = rb < a.nrows && vpad_row(vp, rb, &sb);
  const float* pa = a.src + sa * vp.n[0];
  const float* pb = a.src + sb * vp.n[0];
  float2 v[P::MAXE()];
  sfft_for_inputs<L, false>(t, [&](int slot, int pos) {
    bool okx;
    const int xs = vpad_src(vp, 0, pos, &okx);
    float xa = 0.f, xb = 0.f;
    if (okx) {
      if (oka) xa = pa[xs];
      if (okb) xb = pb[xs];
    }
    v[slot] = make_float2(xa, xb);
  });
  float2* row = lds + c * LP;
  auto addr = [&](int pos) { return XRow<L>::at(pos); };
  sfft_run<L, false, -1, 9>(v, t, row, addr, a.fft.stw);
  sfft_for_outputs<L, false>(t, [&](int slot, int pos) { row[addr(pos)] = v[slot]; });
  __syncthreads();
  float2* da = a.dst + ra * a.dst_row;
  float2* db = a.dst + rb * a.dst_row;
  const bool wa = ra < a.nrows, wb = rb < a.nrows;
  for (int kb = t; kb < a.nbins; kb += T) {
    const float2 Z = row[addr(kb)];
    const float2 Zm = row[addr(kb == 0 ? 0 : L - kb)];
    if (wa) da[kb] = make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
    if (wb) db[kb] = make_float2(0.5f * (Z.y + Zm.y), 0.5f * (Zm.x - Z.x));
  }
}

// kernel rows first (xfwd_body), then the padded image rows
template <int L, int PAIRS>
__global__ void __launch_bounds__(PAIRS * threads_per_transform<L>())
sk_xfwd_vpad(XPass2 a2) {
  __shared__ float2 lds[PAIRS * XRow<L>::LP];
  if ((long)blockIdx.x >= a2.k_blocks)
    xfwd_vpad_body<L, PAIRS>(a2.m, (long)blockIdx.x - a2.k_blocks, lds);
  else
    xfwd_body<L, PAIRS>(a2.k, blockIdx.x, lds);
}

template <int L, int PAIRS>
__global__ void __launch_bounds__(PAIRS * threads_per_transform<L>())
sk_xfwd(XPass2 a2) {
  __shared__ float2 lds[PAIRS * XRow<L>::LP];
  // the kernel's rows take the first k_blocks workgroups so they run
  // alongside the image rows instead of trailing them
  const bool main = (long)blockIdx.x >= a2.k_blocks;
  xfwd_body<L, PAIRS>(main ? a2.m : a2.k, main ? (long)blockIdx.x - a2.k_blocks : (long)blockIdx.x,
                      lds);
}


// ---------------------------------------------------------------------------
// x pass inverse (C2R), two rows per complex FFT
// ---------------------------------------------------------------------------
template <int L, int PAIRS>
__device__ __forceinline__ void xinv_body(const XPassArgs& a, long blk, float2* lds) {
  using P = PlanOf<L, false>;
  constexpr int T = P::T, LP = XRow<L>::LP;
  const int t = threadIdx.x % T, c = threadIdx.x / T;
  const long ra = 2 * (blk * PAIRS + c), rb = ra + 1;
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
  sfft_run<L, false, +1>(v, t, row, addr, a.fft.stw);
  float* oa = a.rdst + ra * a.dst_row;
  float* ob = a.rdst + rb * a.dst_row;
  sfft_for_outputs<L, false>(t, [&](int slot, int pos) {
    if (oka) oa[pos] = v[slot].x;
    if (okb) ob[pos] = v[slot].y;
  });
}

// Same transform, but every half-spectrum bin of the two rows is read from
// HBM once (coalesced) into LDS and the Hermitian mirror is formed there.
template <int L, int PAIRS>
__device__ __forceinline__ void xinv_lds_body(const XPassArgs& a, long blk, float2* lds) {
  using P = PlanOf<L, false>;
  constexpr int T = P::T, LP = XRow<L>::LP;
  const int t = threadIdx.x % T, c = threadIdx.x / T;
  const long ra = 2 * (blk * PAIRS + c), rb = ra + 1;
  const bool oka = ra < a.nrows, okb = rb < a.nrows;
  const float2* pa = a.csrc + ra * a.src_row;
  const float2* pb = a.csrc + rb * a.src_row;
  const int nb = a.nbins;
  float2* row = lds + c * LP;
  auto addr = [&](int pos) { return XRow<L>::at(pos); };
  // Z = A' + iB' over all L positions, built from the nb bins of each row:
  // position kb takes A + iB, its mirror L - kb takes conj(A) + i conj(B)
  for (int kb = t; kb < nb; kb += T) {
    float2 A = make_float2(0.f, 0.f), B = make_float2(0.f, 0.f);
    if (oka) A = pa[kb];
    if (okb) B = pb[kb];
    if (kb == 0 || 2 * kb == L) {
      A.y = 0.f;
      B.y = 0.f;
    }
    // Z[kb] = A + iB and Z[L-kb] = conj(A) + i conj(B)
    row[addr(kb)] = make_float2(A.x - B.y, A.y + B.x);
    if (kb != 0 && 2 * kb != L) row[addr(L - kb)] = make_float2(A.x + B.y, B.x - A.y);
  }
  __syncthreads();
  float2 v[P::MAXE()];
  sfft_for_inputs<L, false>(t, [&](int slot, int pos) { v[slot] = row[addr(pos)]; });
  __syncthreads();
  sfft_run<L, false, +1>(v, t, row, addr, a.fft.stw);
  float* oa = a.rdst + ra * a.dst_row;
  float* ob = a.rdst + rb * a.dst_row;
  sfft_for_outputs<L, false>(t, [&](int slot, int pos) {
    if (oka) oa[pos] = v[slot].x;
    if (okb) ob[pos] = v[slot].y;
  });
}

template <int L, int PAIRS>
__global__ void __launch_bounds__(PAIRS * threads_per_transform<L>())
sk_xinv_lds(XPassArgs a) {
  __shared__ float2 lds[PAIRS * XRow<L>::LP];
  xinv_lds_body<L, PAIRS>(a, blockIdx.x, lds);
}

template <int L, int PAIRS>
__global__ void __launch_bounds__(PAIRS * threads_per_transform<L>())
sk_xinv(XPassArgs a) {
  __shared__ float2 lds[PAIRS * XRow<L>::LP];
  xinv_body<L, PAIRS>(a, blockIdx.x, lds);
}


// ---------------------------------------------------------------------------
// x passes with workgroup-wide ("flat") global access.  In the bodies above
// the T threads of a row pair touch T consecutive elements of each of its
// rows per instruction, so for short rows one wavefront instruction spans up
// to 64/T rows and as many cache lines.  Here the workgroup's rows, which are
// adjacent in memory, move between HBM and an LDS staging area as one run:
// consecutive lanes take consecutive 16-byte (real rows) or 8-byte
// (spectrum bins) pieces.  The staging area aliases the exchange rows.
// ---------------------------------------------------------------------------
template <int L>
struct XStage {
  static constexpr bool V4 = (L % 4) == 0;       // real rows split into float4s
  static constexpr int SL = V4 ? L + 4 : L + 1;  // staging row stride (floats)
  static constexpr int NB = L / 2 + 1;           // half-spectrum bins per row
};

// LDS floats of a flat x-pass workgroup of `pairs` row pairs: the largest of
// the exchange rows, the real staging rows and the spectrum staging rows
template <int L>
constexpr int xflat_lds_floats(int pairs) {
  const int ex = 2 * pairs * XRow<L>::LP;
  const int st = 2 * pairs * XStage<L>::SL;
  const int sc = 4 * pairs * XStage<L>::NB;
  const int m = ex > st ? ex : st;
  return m > sc ? m : sc;
}

// row pairs of a flat x-pass workgroup: about TGT threads, within 1024
// threads and 64 KiB of LDS
template <int L, int TGT>
constexpr int flat_pairs() {
  int p = pairs_for<L, TGT>();
  while (p > 1 && (p * threads_per_transform<L>() > 1024 || xflat_lds_floats<L>(p) * 4 > 65536))
    p /= 2;
  return p;
}

#ifndef FC_BATCH_XTGT
#define FC_BATCH_XTGT 256
#endif
#ifndef FC_BATCH_XFLAT
#define FC_BATCH_XFLAT 1
#endif
// row pairs per workgroup of the batched x passes
template <int L>
constexpr int batch_pairs() {
  return FC_BATCH_XFLAT ? flat_pairs<L, FC_BATCH_XTGT>() : fit_pairs<L, pairs_for<L, FC_BATCH_XTGT>()>();
}

// nr dense rows of L floats at g -> staging rows of stride SL
template <int L, int NT, int MAXR>
__device__ __forceinline__ void xflat_load_rows(const float* __restrict__ g, float* st, int nr) {
  constexpr int SL = XStage<L>::SL;
  if constexpr (XStage<L>::V4) {
    if ((reinterpret_cast<uintptr_t>(g) & 15) == 0) {
      constexpr int Q = L / 4, IT = (MAXR * Q + NT - 1) / NT;
      const int n = nr * Q;
      const float4* g4 = reinterpret_cast<const float4*>(g);
      float4 r[IT];
#pragma unroll
      for (int j = 0; j < IT; ++j) {
        const int i = threadIdx.x + j * NT;
        r[j] = i < n ? g4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
      }
#pragma unroll
      for (int j = 0; j < IT; ++j) {
        const int i = threadIdx.x + j * NT;
        if (i < n) {
          const int row = i / Q;
          *reinterpret_cast<float4*>(st + row * SL + 4 * (i - row * Q)) = r[j];
        }
      }
      return;
    }
  }
  constexpr int IT = (MAXR * L + NT - 1) / NT;
  const int n = nr * L;
  float r[IT];
#pragma unroll
  for (int j = 0; j < IT; ++j) {
    const int i = threadIdx.x + j * NT;
    r[j] = i < n ? g[i] : 0.f;
  }
#pragma unroll
  for (int j = 0; j < IT; ++j) {
    const int i = threadIdx.x + j * NT;
    if (i < n) {
      const int row = i / L;
      st[row * SL + (i - row * L)] = r[j];
    }
  }
}

// staging rows of stride SL -> nr dense rows of L floats at g
template <int L, int NT, int MAXR>
__device__ __forceinline__ void xflat_store_rows(float* __restrict__ g, const float* st, int nr) {
  constexpr int SL = XStage<L>::SL;
  if constexpr (XStage<L>::V4) {
    if ((reinterpret_cast<uintptr_t>(g) & 15) == 0) {
      constexpr int Q = L / 4, IT = (MAXR * Q + NT - 1) / NT;
      const int n = nr * Q;
      float4* g4 = reinterpret_cast<float4*>(g);
#pragma unroll
      for (int j = 0; j < IT; ++j) {
        const int i = threadIdx.x + j * NT;
        if (i < n) {
          const int row = i / Q;
          g4[i] = *reinterpret_cast<const float4*>(st + row * SL + 4 * (i - row * Q));
        }
      }
      return;
    }
  }
  constexpr int IT = (MAXR * L + NT - 1) / NT;
  const int n = nr * L;
#pragma unroll
  for (int j = 0; j < IT; ++j) {
    const int i = threadIdx.x + j * NT;
    if (i < n) {
      const int row = i / L;
      g[i] = st[row * SL + (i - row * L)];
    }
  }
}

// nr spectrum rows (NB bins each, row pitch `pitch`) -> dense staging rows
template <int L, int NT, int MAXR>
__device__ __forceinline__ void xflat_load_bins(const float2* __restrict__ g, long pitch,
                                                float2* sc, int nr) {
  constexpr int NB = XStage<L>::NB, IT = (MAXR * NB + NT - 1) / NT;
  const int n = nr * NB;
  float2 r[IT];
#pragma unroll
  for (int j = 0; j < IT; ++j) {
    const int i = threadIdx.x + j * NT;
    r[j] = make_float2(0.f, 0.f);
    if (i < n) {
      const int row = i / NB;
      r[j] = g[row * pitch + (i - row * NB)];
    }
  }
#pragma unroll
  for (int j = 0; j < IT; ++j) {
    const int i = threadIdx.x + j * NT;
    if (i < n) sc[i] = r[j];
  }
}

// R2C of up to 2*PAIRS dense image rows starting at row 2*blk*PAIRS
template <int L, int PAIRS>
__device__ __forceinline__ void xfwd_flat_body(const XPassArgs& a, long blk, float2* lds) {
  using P = PlanOf<L, false>;
  constexpr int T = P::T, LP = XRow<L>::LP, SL = XStage<L>::SL, NB = XStage<L>::NB;
  constexpr int NT = PAIRS * T, MAXR = 2 * PAIRS;
  const long r0 = 2 * blk * PAIRS;
  if (r0 >= a.nrows) return;  // whole workgroup
  const int nr = a.nrows - r0 < MAXR ? (int)(a.nrows - r0) : MAXR;
  float* st = reinterpret_cast<float*>(lds);
  xflat_load_rows<L, NT, MAXR>(a.src + r0 * L, st, nr);
  __syncthreads();
  const int t = threadIdx.x % T, c = threadIdx.x / T;
  const bool oka = 2 * c < nr, okb = 2 * c + 1 < nr;
  float2 v[P::MAXE()];
  sfft_for_inputs<L, false>(t, [&](int slot, int pos) {
    v[slot] = make_float2(oka ? st[2 * c * SL + pos] : 0.f, okb ? st[(2 * c + 1) * SL + pos] : 0.f);
  });
  __syncthreads();
  float2* row = lds + c * LP;
  auto addr = [&](int pos) { return XRow<L>::at(pos); };
  sfft_run<L, false, -1, 9>(v, t, row, addr, a.fft.stw);
  sfft_for_outputs<L, false>(t, [&](int slot, int pos) { row[addr(pos)] = v[slot]; });
  __syncthreads();
  // Hermitian split, bin-contiguous over the workgroup's rows
  float2* dst = a.dst + r0 * a.dst_row;
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
      dst[r * a.dst_row + kb] = (r & 1) ? make_float2(0.5f * (Z.y + Zm.y), 0.5f * (Zm.x - Z.x))
                                        : make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
    }
  }
}

// C2R of up to 2*PAIRS spectrum rows into dense real rows
template <int L, int PAIRS>
__device__ __forceinline__ void xinv_flat_body(const XPassArgs& a, long blk, float2* lds) {
  using P = PlanOf<L, false>;
  constexpr int T = P::T, LP = XRow<L>::LP, SL = XStage<L>::SL, NB = XStage<L>::NB;
  constexpr int NT = PAIRS * T, MAXR = 2 * PAIRS;
  const long r0 = 2 * blk * PAIRS;
  if (r0 >= a.nrows) return;  // whole workgroup
  const int nr = a.nrows - r0 < MAXR ? (int)(a.nrows - r0) : MAXR;
  xflat_load_bins<L, NT, MAXR>(a.csrc + r0 * a.src_row, a.src_row, lds, nr);
  __syncthreads();
  const int t = threadIdx.x % T, c = threadIdx.x / T;
  const bool oka = 2 * c < nr, okb = 2 * c + 1 < nr;
  const float2* sa = lds + 2 * c * NB;
  const float2* sb = sa + NB;
  float2 v[P::MAXE()];
  // Z = A' + iB' over all L positions; position L - kb takes the conjugates
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
    v[slot] = make_float2(A.x - B.y, A.y + B.x);
  });
  __syncthreads();
  float2* row = lds + c * LP;
  auto addr = [&](int pos) { return XRow<L>::at(pos); };
  sfft_run<L, false, +1>(v, t, row, addr, a.fft.stw);
  // sfft_run ends on a barrier after its last LDS read
  float* st = reinterpret_cast<float*>(lds);
  sfft_for_outputs<L, false>(t, [&](int slot, int pos) {
    st[2 * c * SL + pos] = v[slot].x;
    st[(2 * c + 1) * SL + pos] = v[slot].y;
  });
  __syncthreads();
  xflat_store_rows<L, NT, MAXR>(a.rdst + r0 * L, st, nr);
}

template <int L, int PAIRS>
__global__ void __launch_bounds__(PAIRS * threads_per_transform<L>())
sk_xfwd_flat(XPass2 a2) {
  static_assert(xflat_lds_floats<L>(PAIRS) * 4 <= 65536, "flat x pass LDS");
  __shared__ float2 lds[xflat_lds_floats<L>(PAIRS) / 2];
  if ((long)blockIdx.x >= a2.k_blocks)
    xfwd_flat_body<L, PAIRS>(a2.m, (long)blockIdx.x - a2.k_blocks, lds);
  else
    xfwd_body<L, PAIRS>(a2.k, blockIdx.x, lds);
}

template <int L, int PAIRS>
__global__ void __launch_bounds__(PAIRS * threads_per_transform<L>())
sk_xinv_flat(XPassArgs a) {
  static_assert(xflat_lds_floats<L>(PAIRS) * 4 <= 65536, "flat x pass LDS");
  __shared__ float2 lds[xflat_lds_floats<L>(PAIRS) / 2];
  xinv_flat_body<L, PAIRS>(a, blockIdx.x, lds);
}

// Flat form of the cropped x-C2R of a virtually padded volume (VPad in
// kernels.hpp): the workgroup's spectrum rows are staged as in the flat x
// passes, and each kept row's interior [lo, lo + m) is written with
// coalesced stores.  (A flat forward over the boundary maps, staging whole
// source rows, measured slower than the per-pair body: 258 vs 171 us at
// 600 x 600 x 288.)
template <int L, int PAIRS>
__global__ void __launch_bounds__(PAIRS * threads_per_transform<L>())
sk_xinv_crop_flat(XPassArgs a) {
  using P = PlanOf<L, false>;
  constexpr int T = P::T, LP = XRow<L>::LP, SL = XStage<L>::SL, NB = XStage<L>::NB;
  constexpr int NT = PAIRS * T, MAXR = 2 * PAIRS;
  static_assert(xflat_lds_floats<L>(PAIRS) * 4 <= 65536, "flat x pass LDS");
  __shared__ float2 lds[xflat_lds_floats<L>(PAIRS) / 2];
  const long r0 = 2 * (long)blockIdx.x * PAIRS;
  const int nr = a.nrows - r0 < MAXR ? (int)(a.nrows - r0) : MAXR;
  xflat_load_bins<L, NT, MAXR>(a.csrc + r0 * a.src_row, a.src_row, lds, nr);
  __syncthreads();
  const int t = threadIdx.x % T, c = threadIdx.x / T;
  const bool oka = 2 * c < nr, okb = 2 * c + 1 < nr;
  const float2* sa = lds + 2 * c * NB;
  const float2* sb = sa + NB;
  float2 v[P::MAXE()];
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
    v[slot] = make_float2(A.x - B.y, A.y + B.x);
  });
  __syncthreads();
  float2* row = lds + c * LP;
  auto addr = [&](int pos) { return XRow<L>::at(pos); };
  sfft_run<L, false, +1>(v, t, row, addr, a.fft.stw);
  float* st = reinterpret_cast<float*>(lds);
  sfft_for_outputs<L, false>(t, [&](int slot, int pos) {
    st[2 * c * SL + pos] = v[slot].x;
    st[(2 * c + 1) * SL + pos] = v[slot].y;
  });
  __syncthreads();
  const VPad& vp = a.vp;
  const int m0 = vp.m[0], x0 = vp.lo[0];
  for (int r = 0; r < nr; ++r) {
    long orow = 0;
    if (!vpad_out_row(vp, r0 + r, &orow)) continue;  // uniform across the workgroup
    float* g = a.rdst + orow * m0;
    for (int x = threadIdx.x; x < m0; x += NT) g[x] = st[r * SL + x0 + x];
  }
}

// batched x passes: workgroup i serves tasks[i] = {argument record, row-pair
// group}.  Image rows (dense, unshifted) take the flat bodies; the compact
// kernel's rows the per-pair body.
template <int L, int PAIRS>
__global__ void __launch_bounds__(PAIRS * threads_per_transform<L>())
sb_xfwd(const XPassArgs* __restrict__ args, const int2* __restrict__ tasks) {
  __shared__ float2 lds[xflat_lds_floats<L>(PAIRS) / 2];
  const int2 tk = tasks[blockIdx.x];
  const XPassArgs& a = args[tk.x];
  if (FC_BATCH_XFLAT && a.src_row == L && a.klen == L && a.kcenter == 0)
    xfwd_flat_body<L, PAIRS>(a, tk.y, lds);
  else
    xfwd_body<L, PAIRS>(a, tk.y, lds);
}

template <int L, int PAIRS>
__global__ void __launch_bounds__(PAIRS * threads_per_transform<L>())
sb_xinv(const XPassArgs* __restrict__ args, const int2* __restrict__ tasks) {
  __shared__ float2 lds[xflat_lds_floats<L>(PAIRS) / 2];
  const int2 tk = tasks[blockIdx.x];
  const XPassArgs& a = args[tk.x];
  if (FC_BATCH_XFLAT && a.dst_row == L)
    xinv_flat_body<L, PAIRS>(a, tk.y, lds);
  else
    xinv_lds_body<L, PAIRS>(a, tk.y, lds);
}

// ---------------------------------------------------------------------------
// column pass (y): tiles of TX bins x L positions, numbered by
// column_tile_decode (kernels.hpp)
// ---------------------------------------------------------------------------
template <int L, int TX, bool REV = false>
struct ColTile {
  using P = PlanOf<L, REV>;
  static constexpr int MAXE = P::MAXE();
  const ColArgs* a;
  long base_src, base_dst;
  bool ok;

  __device__ ColTile(const ColArgs& args, long tile, int c) : a(&args) {
    long o;
    int x;
    column_tile_decode(args.nbins, TX, args.nouter, tile, c, &o, &x, &ok);
    base_src = ok ? o * args.src_outer + x : 0;
    base_dst = ok ? o * args.dst_outer + x : 0;
  }
  // Positions are t + (compile-time constant) after unrolling, so the
  // accesses go through a per-thread base at position t and a uniform
  // offset (pos - t) * stride instead of a 64-bit product per element.
  __device__ void load(int t, float2 (&v)[MAXE]) const {
    const float2* src = a->src + base_src;
    const int center = a->center, nvalid = a->nvalid;
    const long es = a->src_elem;
    const bool okk = ok;
    if (center == 0 && nvalid == L) {
      const float2* st = src + (long)t * es;
      sfft_for_inputs<L, REV>(t, [&](int slot, int pos) {
        float2 val = make_float2(0.f, 0.f);
        if (okk) val = st[(long)(pos - t) * es];
        v[slot] = val;
      });
      return;
    }
    // compact support placed circularly (kernel planes): j < nvalid < 2^31 / es
    const int esi = (int)es;
    sfft_for_inputs<L, REV>(t, [&](int slot, int pos) {
      int j = pos + center;
      if (j >= L) j -= L;
      float2 val = make_float2(0.f, 0.f);
      if (okk && j < nvalid) val = src[j * esi];
      v[slot] = val;
    });
  }
  __device__ void store(int t, const float2 (&v)[MAXE]) const {
    if (!ok) return;
    const long es = a->dst_elem;
    float2* dt = a->dst + base_dst + (long)t * es;
    sfft_for_outputs<L, REV>(t, [&](int slot, int pos) { dt[(long)(pos - t) * es] = v[slot]; });
  }
};

// kernel-plane tiles first, then the image's
__device__ __forceinline__ const ColArgs& col2_args(const Col2& a2, long tile, long* local) {
  const long ktiles = a2.ntiles - a2.m_tiles;
  const bool main = tile >= ktiles;
  *local = main ? tile - ktiles : tile;
  return main ? a2.m : a2.k;
}

template <int L, int TX, int DIR, bool REV = false>
__device__ __forceinline__ void col_body(const ColArgs& args, long local, float2* lds) {
  using CT = ColTile<L, TX, REV>;
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  const CT tile(args, local, c);
  float2 v[CT::MAXE];
  tile.load(t, v);
  auto addr = [&](int pos) { return pos * TX + c; };
  sfft_run<L, REV, DIR>(v, t, lds, addr, args.fft.stw);
  tile.store(t, v);
}

// REV runs the radix plan in the opposite stage order
template <int L, int TX, int DIR, bool REV>
__global__ void __launch_bounds__(TX * threads_per_transform<L>())
sk_col(Col2 a2) {
  __shared__ float2 lds[L * TX];
  long local;
  const ColArgs& a = col2_args(a2, blockIdx.x, &local);
  col_body<L, TX, DIR, REV>(a, local, lds);
}

template <int L, int TX, int DIR>
__global__ void __launch_bounds__(TX * threads_per_transform<L>())
sb_col(const ColArgs* __restrict__ args, const int2* __restrict__ tasks) {
  __shared__ float2 lds[L * TX];
  const int2 tk = tasks[blockIdx.x];
  col_body<L, TX, DIR>(args[tk.x], tk.y, lds);
}

template <int L, int TX, int DIR>
__global__ void __launch_bounds__(TX * threads_per_transform<L>())
sk_col_pipe(Col2 a2) {
  using CT = ColTile<L, TX>;
  __shared__ float2 lds[L * TX];
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  auto addr = [&](int pos) { return pos * TX + c; };
  long tile = blockIdx.x;
  float2 v[CT::MAXE], nv[CT::MAXE];
  long local;
  if (tile < a2.ntiles) {
    const ColArgs& a = col2_args(a2, tile, &local);
    CT(a, local, c).load(t, v);
  }
  for (; tile < a2.ntiles; tile += gridDim.x) {
    const long next = tile + gridDim.x;
    if (next < a2.ntiles) {
      long nl;
      const ColArgs& na = col2_args(a2, next, &nl);
      CT(na, nl, c).load(t, nv);
    }
    const ColArgs& a = col2_args(a2, tile, &local);
    const CT cur(a, local, c);
    sfft_run<L, false, DIR>(v, t, lds, addr, a.fft.stw);
    cur.store(t, v);
#pragma unroll
    for (int i = 0; i < CT::MAXE; ++i) v[i] = nv[i];
  }
}

// ---------------------------------------------------------------------------
// fused z pass
// ---------------------------------------------------------------------------
// Kernel spectrum at the thread's last-stage output positions f = b + M*r
// (M = L/R, R = last radix) by direct DFT over the kz taps of the column:
//   K[f] = sum_j k_j W^(f (j - cz)),  W = exp(-2 pi i / L)
// With acc_q = sum_{j = q mod R} k_j W^(b (j - cz)) this is
//   K[b + M r] = W_R^(-r cz) * DFT_R(acc)[r].
template <int L, int MAXE>
__device__ __forceinline__ void kernel_spectrum_direct(float2 (&w)[MAXE], int t, bool ok,
                                                       const float2* __restrict__ kcol, long kplane,
                                                       int kz, int cz,
                                                       const float2* __restrict__ tw) {
  using P = PlanOf<L, false>;
  constexpr int s = P::S - 1;
  constexpr int R = P::R(s), M = P::M(s), NB = P::NB(s);
#pragma unroll
  for (int i = 0; i < NB * R; ++i) w[i] = make_float2(0.f, 0.f);
  // twiddle exponent e_i = b_i * (j - cz) mod L, advanced by b_i per tap
  int e[NB];
  int bb[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    bb[i] = t + i * P::T;
    e[i] = (int)(((long)bb[i] * (L - cz)) % L);
  }
  for (int j0 = 0; j0 < kz; j0 += R) {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int j = j0 + q;
      if (j < kz) {
        float2 kv = make_float2(0.f, 0.f);
        if (ok) kv = kcol[(long)j * kplane];
#pragma unroll
        for (int i = 0; i < NB; ++i) {
          const float2 wt = tw[e[i]];
          w[i * R + q] = make_float2(w[i * R + q].x + kv.x * wt.x - kv.y * wt.y,
                                     w[i * R + q].y + kv.x * wt.y + kv.y * wt.x);
          e[i] += bb[i];
          if (e[i] >= L) e[i] -= L;
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    Dft<R, -1>::run(&w[i * R]);
    // W_R^(-r cz) = conj(tw[M * ((r cz) mod R)])
#pragma unroll
    for (int r = 1; r < R; ++r) w[i * R + r] = cmulc(w[i * R + r], tw[M * ((r * cz) % R)]);
  }
}

template <int L, int TX>
struct ZTile {
  using P = PlanOf<L, false>;
  static constexpr int MAXE = P::MAXE();
  float2* col;
  const float2* kcol;
  bool ok;
  __device__ ZTile(const ZArgs& a, long tile, int c) {
    long y;
    int x;
    column_tile_decode(a.nbins, TX, a.ny, tile, c, &y, &x, &ok);
    if (!ok) {
      y = 0;
      x = 0;
    }
    col = a.s + y * a.s_row + x;
    kcol = a.k + y * a.k_row + x;
  }
  // per-thread base at position t plus a uniform offset (pos - t) * stride
  __device__ void load(const ZArgs& a, int t, float2 (&v)[MAXE]) const {
    const long sp = a.s_plane;
    const bool okk = ok;
    const float2* ct = col + (long)t * sp;
    sfft_for_inputs<L, false>(t, [&](int slot, int pos) {
      float2 val = make_float2(0.f, 0.f);
      if (okk) val = ct[(long)(pos - t) * sp];
      v[slot] = val;
    });
  }
  // SUBSET: honour ZArgs::out_planes (compiled only where a dense operand,
  // the PSF gradient's, can reach the kernel)
  template <bool SUBSET>
  __device__ void store(const ZArgs& a, int t, const float2 (&v)[MAXE]) const {
    if (!ok) return;
    const long sp = a.s_plane;
    if (SUBSET && a.out_planes > 0) {  // rotated subset (see ZArgs)
      const int sh = a.out_shift, np = a.out_planes;
      sfft_for_outputs<L, true>(t, [&](int slot, int pos) {
        int q = pos + sh;
        if (q >= L) q -= L;
        if (q < np) col[(long)q * sp] = v[slot];
      });
      return;
    }
    float2* ct = col + (long)t * sp;
    sfft_for_outputs<L, true>(t, [&](int slot, int pos) { ct[(long)(pos - t) * sp] = v[slot]; });
  }
  // kernel column: taps j < nvalid placed circularly at (j - center) mod L
  __device__ void load_kernel(const ZArgs& a, int t, float2 (&w)[MAXE]) const {
    const long kp = a.k_plane;
    const int center = a.center, nvalid = a.nvalid;
    const bool okk = ok;
    // tap j = pos + center (- L when it wraps): base at tap t + center
    const float2* kt = kcol + (long)(t + center) * kp;
    const float2* kw = kt - (long)L * kp;
    sfft_for_inputs<L, false>(t, [&](int slot, int pos) {
      const bool wrap = pos + center >= L;
      const int j = wrap ? pos + center - L : pos + center;
      float2 val = make_float2(0.f, 0.f);
      if (okk && j < nvalid) val = (wrap ? kw : kt)[(long)(pos - t) * kp];
      w[slot] = val;
    });
  }
};

// product of the forward image spectrum (v) with the kernel spectrum (w),
// conjugated for correlation, times the scale: v * (sc*w.x, +-sc*w.y)
template <int L, int MAXE>
__device__ __forceinline__ void zproduct(float2 (&v)[MAXE], const float2 (&w)[MAXE], int t,
                                         bool cj, float sc) {
  const float si = cj ? -sc : sc;
  sfft_for_outputs<L, false>(t, [&](int slot, int) {
    v[slot] = cmul(v[slot], make_float2(sc * w[slot].x, si * w[slot].y));
  });
}

// the same product with the scale already in w (the tap path scales its
// taps once as it loads them): v * (w.x, +-w.y)
template <int L, int MAXE>
__device__ __forceinline__ void zproduct_prescaled(float2 (&v)[MAXE], const float2 (&w)[MAXE], int t,
                                                   bool cj) {
  const float si = cj ? -1.f : 1.f;
  sfft_for_outputs<L, false>(t, [&](int slot, int) {
    v[slot] = cmul(v[slot], make_float2(w[slot].x, si * w[slot].y));
  });
}

// taps per column the two-stage fast path can hold (its stage-1 radix)
template <int L>
constexpr int ktap_slots() {
  return PlanOf<L, false>::S == 2 ? PlanOf<L, false>::R(1) : 1;
}

// Two-stage plan (L = R0*R1) with the kz taps within one stage-0 stride
// (nvalid <= R1): stage-0 butterfly r of the kernel column (positions
// r + rho*R1) meets at most one tap, rho = 0 (tap r + center) or
// rho = R0 - 1 (tap r + center - R1), so its output k is that tap times 1 or
// exp(+2 pi i k / R0).  Output k = b of butterfly r is the kernel's stage-1
// input slot (i, r) of the thread owning stage-1 butterfly b = t + i*T.
// Each thread loads at most ceil(kz / T) taps into `ktaps`; the image
// transform's stage-0 exchange barrier publishes them, and the stage-1
// inputs are read back by broadcast.  Compared with transforming the
// kernel column like the image's, this drops its stage-0 butterflies, its
// per-slot predicated loads and one LDS exchange (two per tile remain).
template <int L, int TX>
__device__ __forceinline__ void zfused_sparse_body(const ZArgs& a, const ZTile<L, TX>& tile, int t,
                                                   int c, float2 (&v)[ZTile<L, TX>::MAXE],
                                                   float2* lds, float2* ktaps) {
  using P = PlanOf<L, false>;
  constexpr int R1 = P::R(1), M1 = P::M(1), NB1 = P::NB(1), T = P::T;
  constexpr int MAXE = ZTile<L, TX>::MAXE;
  auto addr = [&](int pos) { return pos * TX + c; };
  const int center = a.center, nvalid = a.nvalid;
  for (int j = t; j < nvalid; j += T)
    ktaps[j * TX + c] = tile.ok ? cscale(tile.kcol[(long)j * a.k_plane], a.scale) : make_float2(0.f, 0.f);
  // exp(+2 pi i b / R0) = conj(tw[b * R1]), tw[j] = exp(-2 pi i j / L)
  float2 e[NB1];
#pragma unroll
  for (int i = 0; i < NB1; ++i) {
    const int b = t + i * T;
    e[i] = (M1 % T) != 0 && b >= M1 ? make_float2(1.f, 0.f) : cconj(a.fft.tw[b * R1]);
  }
  // image: stage 0 and its exchange; the barrier also publishes the taps
  sfft_compute<L, false, -1, 0>(v, t, a.fft.stw);
  sfft_store_lds<L, false, 0>(v, t, lds, addr);
  StageTw<L, false, 1> tw1;
  sfft_load_tw<L, false, 1>(tw1, t, a.fft.stw);
  __syncthreads();
  sfft_load_lds<L, false, 1>(v, t, lds, addr);
  // every thread is done with the exchange tile before the inverse reuses it
  // (the taps live in their own array)
  __syncthreads();
  sfft_twiddle<L, false, -1, 1>(v, t, tw1);
  sfft_butterflies<L, false, -1, 1>(v, t);
  // the kernel column's stage-1 inputs, built only now to keep them out of
  // the image transform's register peak; same butterflies, same twiddles
  float2 w[MAXE];
#pragma unroll
  for (int i = 0; i < NB1; ++i) {
#pragma unroll
    for (int r = 0; r < R1; ++r) {
      const int j0 = r + center, j1 = r + center - R1;
      float2 val = make_float2(0.f, 0.f);
      if (j0 < nvalid)
        val = ktaps[j0 * TX + c];
      else if (j1 >= 0)
        val = cmul(ktaps[j1 * TX + c], e[i]);
      w[i * R1 + r] = val;
    }
  }
  sfft_twiddle<L, false, -1, 1>(w, t, tw1);
  sfft_butterflies<L, false, -1, 1>(w, t);
  zproduct_prescaled<L>(v, w, t, a.conj != 0);
  sfft_run<L, true, +1>(v, t, lds, addr, a.fft.stw);
  tile.template store<false>(a, t, v);
}

// whether the two-stage tap path applies to this launch
template <int L>
__host__ __device__ constexpr bool zfused_sparse_ok(int nvalid) {
  return PlanOf<L, false>::S == 2 && nvalid <= PlanOf<L, false>::R(1);
}

// Fused z pass of one tile.  PATH 0: kernel column transformed like the
// image's; 1: the two-stage tap path (caller guarantees zfused_sparse_ok);
// 2: chosen per tile at run time (batched task lists mix kernels).  Paths 0
// and 1 are separate instantiations: compiled into one kernel their register
// peaks add up (~144 against ~105 VGPRs each).
template <int L, int TX, int PATH>
__device__ __forceinline__ void zfused_fft_body(const ZArgs& a, long tile_index, float2* lds,
                                                float2* ktaps) {
  using ZT = ZTile<L, TX>;
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  const ZT tile(a, tile_index, c);
  auto addr = [&](int pos) { return pos * TX + c; };
  float2 v[ZT::MAXE];
  tile.load(a, t, v);
  if constexpr (PlanOf<L, false>::S == 2 && PATH != 0) {
    if (PATH == 1 || a.nvalid <= PlanOf<L, false>::R(1)) {
      zfused_sparse_body<L, TX>(a, tile, t, c, v, lds, ktaps);
      return;
    }
  }
  (void)ktaps;
  float2 w[ZT::MAXE];
  tile.load_kernel(a, t, w);
  // every sfft_run ends with a barrier after its last LDS read, so the
  // next transform may reuse the tile
  sfft_run<L, false, -1>(v, t, lds, addr, a.fft.stw);
  sfft_run<L, false, -1>(w, t, lds, addr, a.fft.stw);
  zproduct<L>(v, w, t, a.conj != 0, a.scale);
  sfft_run<L, true, +1>(v, t, lds, addr, a.fft.stw);
  tile.template store<PATH == 0>(a, t, v);
}

template <int L, int TX, int PATH>
__global__ void __launch_bounds__(TX * threads_per_transform<L>())
sk_zfused_fft(ZArgs a) {
  __shared__ float2 lds[L * TX];
  __shared__ float2 ktaps[TX * (PATH == 0 ? 1 : ktap_slots<L>())];
  zfused_fft_body<L, TX, PATH>(a, blockIdx.x, lds, ktaps);
}

template <int L, int TX, int PATH>
__global__ void __launch_bounds__(TX * threads_per_transform<L>())
sb_zfused(const ZArgs* __restrict__ args, const int2* __restrict__ tasks) {
  __shared__ float2 lds[L * TX];
  __shared__ float2 ktaps[TX * (PATH == 0 ? 1 : ktap_slots<L>())];
  const int2 tk = tasks[blockIdx.x];
  zfused_fft_body<L, TX, PATH>(args[tk.x], tk.y, lds, ktaps);
}

template <int L, int TX>
__global__ void __launch_bounds__(TX * threads_per_transform<L>())
sk_zfused_dft(ZArgs a) {
  using ZT = ZTile<L, TX>;
  __shared__ float2 lds[L * TX];
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  const ZT tile(a, blockIdx.x, c);
  auto addr = [&](int pos) { return pos * TX + c; };
  float2 v[ZT::MAXE];
  tile.load(a, t, v);
  sfft_run<L, false, -1>(v, t, lds, addr, a.fft.stw);
  float2 w[ZT::MAXE];
  kernel_spectrum_direct<L>(w, t, tile.ok, tile.kcol, a.k_plane, a.nvalid, a.center, a.fft.tw);
  zproduct<L>(v, w, t, a.conj != 0, a.scale);
  sfft_run<L, true, +1>(v, t, lds, addr, a.fft.stw);
  tile.template store<false>(a, t, v);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// x-pass launchers take the target thread count per workgroup (TGT)
template <int L, int TGT>
bool sk_launch_xfwd(const XPass2& a, hipStream_t st) {
  constexpr int P = fit_pairs<L, pairs_for<L, TGT>()>();
  XPass2 b = a;
  b.m_blocks = ((a.m.nrows + 1) / 2 + P - 1) / P;
  b.k_blocks = ((a.k.nrows + 1) / 2 + P - 1) / P;
  hipLaunchKernelGGL((sk_xfwd<L, P>), dim3((unsigned)(b.m_blocks + b.k_blocks)),
                     dim3(P * threads_per_transform<L>()), 0, st, b);
  return true;
}

template <int L, int TGT>
bool sk_launch_xinv(const XPassArgs& a, hipStream_t st) {
  constexpr int P = fit_pairs<L, pairs_for<L, TGT>()>();
  const long blocks = ((a.nrows + 1) / 2 + P - 1) / P;
  hipLaunchKernelGGL((sk_xinv<L, P>), dim3((unsigned)blocks), dim3(P * threads_per_transform<L>()),
                     0, st, a);
  return true;
}

template <int L, int TGT>
bool sk_launch_xfwd_vpad(const XPass2& a, hipStream_t st) {
  constexpr int P = fit_pairs<L, pairs_for<L, TGT>()>();
  XPass2 b = a;
  b.m_blocks = ((a.m.nrows + 1) / 2 + P - 1) / P;
  b.k_blocks = ((a.k.nrows + 1) / 2 + P - 1) / P;
  hipLaunchKernelGGL((sk_xfwd_vpad<L, P>), dim3((unsigned)(b.m_blocks + b.k_blocks)),
                     dim3(P * threads_per_transform<L>()), 0, st, b);
  return true;
}

template <int L, int TGT>
bool sk_launch_xinv_crop_flat(const XPassArgs& a, hipStream_t st) {
  constexpr int P = flat_pairs<L, TGT>();
  const long blocks = ((a.nrows + 1) / 2 + P - 1) / P;
  hipLaunchKernelGGL((sk_xinv_crop_flat<L, P>), dim3((unsigned)blocks),
                     dim3(P * threads_per_transform<L>()), 0, st, a);
  return true;
}

template <int L, int TGT>
bool sk_launch_xinv_lds(const XPassArgs& a, hipStream_t st) {
  constexpr int P = fit_pairs<L, pairs_for<L, TGT>()>();
  const long blocks = ((a.nrows + 1) / 2 + P - 1) / P;
  hipLaunchKernelGGL((sk_xinv_lds<L, P>), dim3((unsigned)blocks),
                     dim3(P * threads_per_transform<L>()), 0, st, a);
  return true;
}

// flat variants need dense real rows (row stride L)
template <int L, int TGT>
bool sk_launch_xfwd_flat(const XPass2& a, hipStream_t st) {
  if (a.m.nrows > 0 && a.m.src_row != L) return false;
  constexpr int P = flat_pairs<L, TGT>();
  XPass2 b = a;
  b.m_blocks = ((a.m.nrows + 1) / 2 + P - 1) / P;
  b.k_blocks = ((a.k.nrows + 1) / 2 + P - 1) / P;
  hipLaunchKernelGGL((sk_xfwd_flat<L, P>), dim3((unsigned)(b.m_blocks + b.k_blocks)),
                     dim3(P * threads_per_transform<L>()), 0, st, b);
  return true;
}

template <int L, int TGT>
bool sk_launch_xinv_flat(const XPassArgs& a, hipStream_t st) {
  if (a.dst_row != L) return false;
  constexpr int P = flat_pairs<L, TGT>();
  const long blocks = ((a.nrows + 1) / 2 + P - 1) / P;
  hipLaunchKernelGGL((sk_xinv_flat<L, P>), dim3((unsigned)blocks),
                     dim3(P * threads_per_transform<L>()), 0, st, a);
  return true;
}

template <int L, int TX>
Col2 col2_tiles(const Col2& a) {
  Col2 b = a;
  b.ntx = (a.m.nbins + TX - 1) / TX;
  b.m.nouter = a.m_outer;
  b.k.nouter = a.k_outer;
  b.m_tiles = column_tiles(a.m.nbins, TX, a.m_outer);
  b.ntiles = b.m_tiles + (a.k_outer > 0 ? column_tiles(a.k.nbins, TX, a.k_outer) : 0);
  return b;
}

template <int L, int WANT, bool REV = false>
bool sk_launch_col(const Col2& a, int dir, hipStream_t st) {
  if (REV && PlanOf<L, false>::S < 2) return false;  // one stage: same as forward order
  constexpr int TX = fit_tile<L, WANT>();
  const Col2 b = col2_tiles<L, TX>(a);
  const dim3 block(TX * threads_per_transform<L>());
  if (dir < 0)
    hipLaunchKernelGGL((sk_col<L, TX, -1, REV>), dim3((unsigned)b.ntiles), block, 0, st, b);
  else
    hipLaunchKernelGGL((sk_col<L, TX, +1, REV>), dim3((unsigned)b.ntiles), block, 0, st, b);
  return true;
}

template <int L, int WANT>
bool sk_launch_col_pipe(const Col2& a, int dir, hipStream_t st) {
  constexpr int TX = fit_tile<L, WANT>();
  constexpr int threads = TX * threads_per_transform<L>();
  const Col2 b = col2_tiles<L, TX>(a);
  const void* fn = dir < 0 ? (const void*)sk_col_pipe<L, TX, -1> : (const void*)sk_col_pipe<L, TX, +1>;
  long grid = persistent_blocks(fn, threads);
  if (grid > b.ntiles) grid = b.ntiles;
  if (grid < 1) grid = 1;
  if (dir < 0)
    hipLaunchKernelGGL((sk_col_pipe<L, TX, -1>), dim3((unsigned)grid), dim3(threads), 0, st, b);
  else
    hipLaunchKernelGGL((sk_col_pipe<L, TX, +1>), dim3((unsigned)grid), dim3(threads), 0, st, b);
  return true;
}

template <int L, int WANT>
bool sk_launch_zfused_fft(const ZArgs& a, long ny, hipStream_t st) {
  constexpr int TX = fit_tile<L, WANT>();
  ZArgs b = a;
  b.ny = ny;
  const dim3 grid((unsigned)column_tiles(a.nbins, TX, ny)), block(TX * threads_per_transform<L>());
  if constexpr (PlanOf<L, false>::S == 2) {
    if (zfused_sparse_ok<L>(a.nvalid)) {
      hipLaunchKernelGGL((sk_zfused_fft<L, TX, 1>), grid, block, 0, st, b);
      return true;
    }
  }
  hipLaunchKernelGGL((sk_zfused_fft<L, TX, 0>), grid, block, 0, st, b);
  return true;
}

// direct kernel DFT costs kz complex MACs per output; use it while that is
// cheaper than a transform of the mostly-zero kernel column
template <int L>
constexpr int dft_kz_limit() {
  return L / 4 > 64 ? 64 : L / 4;
}

template <int L, int WANT>
bool sk_launch_zfused_dft(const ZArgs& a, long ny, hipStream_t st) {
  if (a.nvalid > dft_kz_limit<L>()) return false;
  constexpr int TX = fit_tile<L, WANT>();
  ZArgs b = a;
  b.ny = ny;
  hipLaunchKernelGGL((sk_zfused_dft<L, TX>), dim3((unsigned)column_tiles(a.nbins, TX, ny)),
                     dim3(TX * threads_per_transform<L>()), 0, st, b);
  return true;
}

// batched launches over device task tables (default variants)
template <int L>
void sb_launch_xfwd(const XPassArgs* args, const int2* tasks, int ntasks, hipStream_t st) {
  constexpr int P = batch_pairs<L>();
  hipLaunchKernelGGL((sb_xfwd<L, P>), dim3((unsigned)ntasks), dim3(P * threads_per_transform<L>()),
                     0, st, args, tasks);
}
template <int L>
void sb_launch_xinv(const XPassArgs* args, const int2* tasks, int ntasks, hipStream_t st) {
  constexpr int P = batch_pairs<L>();
  hipLaunchKernelGGL((sb_xinv<L, P>), dim3((unsigned)ntasks), dim3(P * threads_per_transform<L>()),
                     0, st, args, tasks);
}
template <int L>
void sb_launch_col(const ColArgs* args, const int2* tasks, int ntasks, int dir, hipStream_t st) {
  constexpr int TX = fit_tile<L, 16>();
  const dim3 block(TX * threads_per_transform<L>());
  if (dir < 0)
    hipLaunchKernelGGL((sb_col<L, TX, -1>), dim3((unsigned)ntasks), block, 0, st, args, tasks);
  else
    hipLaunchKernelGGL((sb_col<L, TX, +1>), dim3((unsigned)ntasks), block, 0, st, args, tasks);
}
template <int L>
void sb_launch_zfused(const ZArgs* args, const int2* tasks, int ntasks, hipStream_t st) {
  constexpr int TX = fit_tile<L, 16>();
  hipLaunchKernelGGL((sb_zfused<L, TX, 0>), dim3((unsigned)ntasks),
                     dim3(TX * threads_per_transform<L>()), 0, st, args, tasks);
}
template <int L>
void sb_launch_zfused_sparse(const ZArgs* args, const int2* tasks, int ntasks, hipStream_t st) {
  constexpr int TX = fit_tile<L, 16>();
  if constexpr (PlanOf<L, false>::S == 2)
    hipLaunchKernelGGL((sb_zfused<L, TX, 1>), dim3((unsigned)ntasks),
                       dim3(TX * threads_per_transform<L>()), 0, st, args, tasks);
}

template <int L>
StaticEntry make_static_entry() {
  StaticEntry e;
  e.xfwd = {{"w256", &sk_launch_xfwd<L, 256>},
            {"w128", &sk_launch_xfwd<L, 128>},
            {"w64", &sk_launch_xfwd<L, 64>},
            {"w256-flat", &sk_launch_xfwd_flat<L, 256>},
            {"w128-flat", &sk_launch_xfwd_flat<L, 128>}};
  e.xinv = {{"w256", &sk_launch_xinv<L, 256>},
            {"w128", &sk_launch_xinv<L, 128>},
            {"w64", &sk_launch_xinv<L, 64>},
            {"w256-ldsmirror", &sk_launch_xinv_lds<L, 256>},
            {"w256-flat", &sk_launch_xinv_flat<L, 256>},
            {"w128-flat", &sk_launch_xinv_flat<L, 128>}};
  e.col = {{"t16", &sk_launch_col<L, 16>, fit_tile<L, 16>()},
           {"t16-pipe", &sk_launch_col_pipe<L, 16>},
           {"t8-pipe", &sk_launch_col_pipe<L, 8>},
           {"t16-revradix", &sk_launch_col<L, 16, true>},
           {"t8", &sk_launch_col<L, 8>, fit_tile<L, 8>()},
           {"t32", &sk_launch_col<L, 32>, fit_tile<L, 32>()},
           {"t64", &sk_launch_col<L, 64>, fit_tile<L, 64>()}};
  e.zfused = {{"t16-fft", &sk_launch_zfused_fft<L, 16>},
              {"t16-dft", &sk_launch_zfused_dft<L, 16>},
              {"t8-fft", &sk_launch_zfused_fft<L, 8>},
              {"t32-fft", &sk_launch_zfused_fft<L, 32>},
              {"t64-fft", &sk_launch_zfused_fft<L, 64>}};
  // defaults (measured over (512, 1024, L) volumes): the flat x inverse at
  // every length; the flat x forward where rows are short or not a whole
  // number of float4s, the per-pair forward elsewhere
  auto index_of = [](const auto& vs, const char* name) {
    for (size_t i = 0; i < vs.size(); ++i)
      if (__builtin_strcmp(vs[i].name, name) == 0) return (int)i;
    return 0;
  };
  e.selected[kSxInv] = index_of(e.xinv, "w256-flat");
  if (L < 128 || L % 4 != 0) e.selected[kSxFwd] = index_of(e.xfwd, "w256-flat");
  e.xfwd_vpad = &sk_launch_xfwd_vpad<L, 64>;
  e.xinv_crop = &sk_launch_xinv_crop_flat<L, 256>;  // 171 vs 193 us at 600 x 600 x 288
  e.batch.pairs = batch_pairs<L>();
  e.batch.tx = fit_tile<L, 16>();
  e.batch.xfwd = &sb_launch_xfwd<L>;
  e.batch.xinv = &sb_launch_xinv<L>;
  e.batch.col = &sb_launch_col<L>;
  e.batch.zfused = &sb_launch_zfused<L>;
  if constexpr (PlanOf<L, false>::S == 2) {
    e.batch.zfused_sparse = &sb_launch_zfused_sparse<L>;
    e.batch.sparse_kz = PlanOf<L, false>::R(1);
  }
  return e;
}

}  // namespace fcamd
