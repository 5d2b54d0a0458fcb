// fft_static.hpp -- compile-time-length Stockham FFT held in registers.
//
// A transform of length L is served by T threads (L / Rmax for the classic
// plan, Rmax = largest radix; chosen by a cost model for composite plans).  In stage s (radix R, Ns = product of the earlier
// radices, M = L / R) thread t owns butterflies b = t + i*T (i < NB,
// predicated b < M); butterfly b reads positions b + r*M and writes
// positions (b / Ns)*Ns*R + b % Ns + r*Ns (Stockham auto-sort, so the result
// is in natural order).  Stage 0 inputs come straight from the caller's
// loader and last-stage outputs stay in registers at positions
// b + r*(L / R_last); only the exchanges between stages go through LDS.
//
// A plan run in reverse order (REV) starts with the radix the forward plan
// ended with, so the same thread already holds exactly the positions the
// reversed plan's first stage needs: a forward transform, a pointwise
// product and an inverse transform chain with no LDS exchange in between.
#pragma once

#include "fc_internal.hpp"
#include "fft_radix.hpp"

namespace fcamd {

struct StaticPlanData {
  int n = 0;
  int r[16] = {};
  int t = 0;  // threads per transform
};

#ifndef FC_POW2_MAXBITS
#define FC_POW2_MAXBITS 4
#endif

// Classic plan for a 13-smooth length: the power-of-two part in the fewest
// stages of radix <= 16 with the bits spread evenly (256 = 16*16,
// 512 = 8*8*8, 1024 = 16*8*8, 128 = 16*8), then 3s, 5s, 7s, 11s and 13s; one thread
// per largest-radix butterfly.
constexpr StaticPlanData classic_plan(int L) {
  StaticPlanData p{};
  int a = 0;
  int n = L;
  while (n % 2 == 0) { n /= 2; ++a; }
  if (a > 0) {
    const int st = (a + FC_POW2_MAXBITS - 1) / FC_POW2_MAXBITS;
    const int lo = a / st, extra = a % st;
    for (int i = 0; i < st; ++i) p.r[p.n++] = 1 << (lo + (i < extra ? 1 : 0));
  }
  while (n % 3 == 0) { p.r[p.n++] = 3; n /= 3; }
  while (n % 5 == 0) { p.r[p.n++] = 5; n /= 5; }
  while (n % 7 == 0) { p.r[p.n++] = 7; n /= 7; }
  while (n % 11 == 0) { p.r[p.n++] = 11; n /= 11; }
  while (n % 13 == 0) { p.r[p.n++] = 13; n /= 13; }
  if (n != 1) {
    p.n = -1;  // not 13-smooth: no static plan
    return p;
  }
  int rmax = 1;
  for (int i = 0; i < p.n; ++i) rmax = p.r[i] > rmax ? p.r[i] : rmax;
  p.t = L / rmax;
  return p;
}

// Composite plan: the fewest stages with radices from {2..16} including
// the composite butterflies 6, 9, 10, 12, 14, 15 (240 = 16*15,
// 210 = 15*14, 500 = 10*10*5), and the thread count per transform T (one of
// the stages' butterfly counts M = L / R) chosen with the radix multiset to
// minimise a per-lane cost model: each stage costs T * ceil(M / T)
// butterflies (idle lanes included) of an estimated flop count plus
// twiddle and LDS-exchange work.  At most 32 values per thread are held in
// registers.
namespace plan_detail {
constexpr int kRadix[15] = {16, 15, 14, 13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2};
constexpr int butterfly_flops(int r) {
  switch (r) {
    case 2: return 4;
    case 3: return 16;
    case 4: return 16;
    case 5: return 40;
    case 6: return 56;
    case 7: return 72;
    case 8: return 52;
    case 9: return 120;
    case 10: return 124;
    case 11: return 240;  // direct pair form: 2 * 5 * 5 complex multiply-adds
    case 13: return 340;
    case 12: return 148;
    case 14: return 208;
    case 15: return 248;
    case 16: return 144;
    default: return 1 << 20;
  }
}
struct Best {
  StaticPlanData plan{};
  long cost = -1;
};
constexpr void consider(const StaticPlanData& cur, int L, Best& best) {
  // candidate thread counts: one butterfly per thread in some stage
  for (int c = 0; c < cur.n; ++c) {
    const int T = L / cur.r[c];
    if (T > 256) continue;
    long cost = 0;
    bool ok = true;
    for (int s = 0; s < cur.n; ++s) {
      const int R = cur.r[s], M = L / R, nb = (M + T - 1) / T;
      if (nb * R > 32) ok = false;
      const int w = butterfly_flops(R) + (s > 0 ? 6 * (R - 1) + 4 * R : 0);
      cost += (long)T * nb * w;
    }
    if (ok && (best.cost < 0 || cost < best.cost)) {
      best.cost = cost;
      best.plan = cur;
      best.plan.t = T;
    }
  }
}
constexpr void search(int rem, int first, StaticPlanData& cur, int S, int L, Best& best) {
  if (cur.n == S) {
    if (rem == 1) consider(cur, L, best);
    return;
  }
  for (int i = first; i < 15; ++i) {
    const int R = kRadix[i];
    if (rem % R != 0) continue;
    cur.r[cur.n++] = R;
    search(rem / R, i, cur, S, L, best);
    --cur.n;
  }
}
}  // namespace plan_detail

constexpr StaticPlanData composite_plan(int L) {
  const StaticPlanData c = classic_plan(L);
  if (c.n <= 0) return c;
  for (int S = 1; S <= c.n; ++S) {
    plan_detail::Best best{};
    StaticPlanData cur{};
    plan_detail::search(L, 0, cur, S, L, best);
    if (best.cost >= 0) return best.plan;
  }
  return c;
}

#ifndef FC_PLAN_COMPOSITE
#define FC_PLAN_COMPOSITE 1
#endif

// Plan family used for length L: the composite plan where it needs fewer
// stages (one LDS exchange less per stage saved) than the classic plan.
// Measured on MI355X over all five passes (L x 256 x 256 images): faster
// for 105, 120, 144, 160, 192, 210, 224 and 240; slower for 96, whose
// composite plan (16*6) leaves 6 threads per transform.
constexpr bool uses_composite_plan(int L) {
  return FC_PLAN_COMPOSITE != 0 && L != 96 && composite_plan(L).n < classic_plan(L).n;
}

constexpr StaticPlanData static_plan(int L) {
  return uses_composite_plan(L) ? composite_plan(L) : classic_plan(L);
}

// Per-stage twiddle tables of a static plan, stored for the forward stage
// order and then for the reversed order: stage s holds Ns(s) rows of
// even(R) float2, row k = exp(-2 pi i r k / (Ns R)) for r = 0..R-1, so a
// thread fetches its butterfly's twiddles with R/2 aligned 16-byte loads.
constexpr int even_up(int r) { return (r + 1) / 2 * 2; }

constexpr int static_tw_size(const StaticPlanData& p, bool rev) {
  int total = 0, ns = 1;
  for (int i = 0; i < p.n; ++i) {
    const int r = rev ? p.r[p.n - 1 - i] : p.r[i];
    total += ns * even_up(r);
    ns *= r;
  }
  return total;
}

constexpr int plan_rmax(const StaticPlanData& p) {
  int m = 1;
  for (int i = 0; i < p.n; ++i) m = p.r[i] > m ? p.r[i] : m;
  return m;
}

template <int L, bool REV>
struct PlanOf {
  static constexpr StaticPlanData data = static_plan(L);
  static_assert(data.n > 0, "length without a static plan");
  static constexpr int S = data.n;
  static constexpr int RMAX = plan_rmax(data);
  static constexpr int T = data.t;
  static constexpr int R(int s) { return REV ? data.r[S - 1 - s] : data.r[s]; }
  static constexpr int Ns(int s) {
    int v = 1;
    for (int i = 0; i < s; ++i) v *= R(i);
    return v;
  }
  static constexpr int M(int s) { return L / R(s); }
  static constexpr int NB(int s) { return (M(s) + T - 1) / T; }
  static constexpr int E(int s) { return NB(s) * R(s); }
  static constexpr int MAXE() {
    int m = 0;
    for (int s = 0; s < S; ++s) m = E(s) > m ? E(s) : m;
    return m;
  }
  // offset (float2 units) of stage s's table in the combined table
  static constexpr int TWOFF(int s) {
    int off = REV ? static_tw_size(data, false) : 0;
    for (int i = 0; i < s; ++i) off += Ns(i) * even_up(R(i));
    return off;
  }
};

// One stage's butterflies on the thread's registers (inputs already loaded
// at positions b + r*M, outputs left in place).
template <int L, bool REV, int DIR, int s, int MAXE>
__device__ __forceinline__ void sfft_compute(float2 (&v)[MAXE], int t, const float2* __restrict__ stw) {
  using P = PlanOf<L, REV>;
  constexpr int R = P::R(s), Ns = P::Ns(s), M = P::M(s), NB = P::NB(s);
  constexpr int RP = even_up(R);
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = t + i * P::T;
    if ((M % P::T) != 0 && b >= M) continue;
    float2* x = &v[i * R];
    if (Ns > 1) {
      const int k = b % Ns;
      const float4* q = reinterpret_cast<const float4*>(stw + P::TWOFF(s) + k * RP);
      float2 w[RP];
#pragma unroll
      for (int r2 = 0; r2 < RP / 2; ++r2) {
        const float4 p = q[r2];
        w[2 * r2] = make_float2(p.x, p.y);
        w[2 * r2 + 1] = make_float2(p.z, p.w);
      }
#pragma unroll
      for (int r = 1; r < R; ++r) x[r] = DIR < 0 ? cmul(x[r], w[r]) : cmulc(x[r], w[r]);
    }
    Dft<R, DIR>::run(x);
  }
}

// Writes stage s outputs to LDS through `addr(pos)`.
template <int L, bool REV, int s, int MAXE, class Addr>
__device__ __forceinline__ void sfft_store_lds(const float2 (&v)[MAXE], int t, float2* lds, Addr addr) {
  using P = PlanOf<L, REV>;
  constexpr int R = P::R(s), Ns = P::Ns(s), M = P::M(s), NB = P::NB(s);
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = t + i * P::T;
    if ((M % P::T) != 0 && b >= M) continue;
    const int k = b % Ns;
    const int base = (b - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) lds[addr(base + r * Ns)] = v[i * R + r];
  }
}

// Reads stage s inputs from LDS.
template <int L, bool REV, int s, int MAXE, class Addr>
__device__ __forceinline__ void sfft_load_lds(float2 (&v)[MAXE], int t, const float2* lds, Addr addr) {
  using P = PlanOf<L, REV>;
  constexpr int R = P::R(s), M = P::M(s), NB = P::NB(s);
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = t + i * P::T;
    if ((M % P::T) != 0 && b >= M) continue;
#pragma unroll
    for (int r = 0; r < R; ++r) v[i * R + r] = lds[addr(b + r * M)];
  }
}

// Stage-0 input positions of the thread: calls f(slot, pos) for every valid
// input (slot indexes the register array).
template <int L, bool REV, class F>
__device__ __forceinline__ void sfft_for_inputs(int t, F f) {
  using P = PlanOf<L, REV>;
  constexpr int R = P::R(0), M = P::M(0), NB = P::NB(0);
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = t + i * P::T;
    if ((M % P::T) != 0 && b >= M) continue;
#pragma unroll
    for (int r = 0; r < R; ++r) f(i * R + r, b + r * M);
  }
}

// Last-stage output positions of the thread: f(slot, pos).
template <int L, bool REV, class F>
__device__ __forceinline__ void sfft_for_outputs(int t, F f) {
  using P = PlanOf<L, REV>;
  constexpr int s = P::S - 1;
  constexpr int R = P::R(s), M = P::M(s), NB = P::NB(s);
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = t + i * P::T;
    if ((M % P::T) != 0 && b >= M) continue;
#pragma unroll
    for (int r = 0; r < R; ++r) f(i * R + r, b + r * M);
  }
}

#ifndef FC_TW_COMPACT
#define FC_TW_COMPACT 1
#endif
#ifndef FC_TW_COMPACT_MIN
#define FC_TW_COMPACT_MIN 8
#endif
// Stage s twiddle rows of the thread's butterflies, loaded ahead of use.
// Stages of radix 8 and above keep bases instead of the row's R - 1
// entries: W^k, W^2k, W^3k (q[0].xy, q[0].zw, q[1].xy), W^4k (q[1].zw) and,
// from radix 9, W^8k and W^12k (q[2]); entry r = r1 + 4*r2 is
// W^(r1 k) * W^(4 r2 k).  That is one product more for the entries with
// r1, r2 > 0, but fewer loads and registers (12 instead of 32 at radix 16):
// one stage's twiddles are live beside the image's and the kernel column's
// values at the fused z pass's register peak, and their loads sit between
// the two barriers of every exchange.  Measured at 512 x 512 x 256: the z
// pass 124 -> 115 us.
template <int L, bool REV, int s, int CM = FC_TW_COMPACT_MIN>
struct StageTw {
  using P = PlanOf<L, REV>;
  static constexpr int R = P::R(s), Ns = P::Ns(s), NB = P::NB(s), RP = even_up(R);
  static constexpr int CB = R >= 9 && R >= CM ? 3 : (R >= CM ? 2 : 0);
  static constexpr bool COMPACT = FC_TW_COMPACT && CB > 0;
  // odd radices below that keep entries 1..R-1 only (the row's entry 0 is
  // 1 and its last is padding): radix 3 in one float4 instead of two
  static constexpr bool SHIFT = FC_TW_COMPACT && !COMPACT && (R & 1) && R >= 3;
  static constexpr int QB = COMPACT ? CB : (SHIFT ? (R - 1) / 2 : RP / 2);  // float4 per butterfly
  static constexpr int Q = Ns > 1 ? NB * QB : 1;
  float4 q[Q];
  // twiddle of entry r (0 < r < R) of butterfly slot i
  __device__ __forceinline__ float2 get(int i, int r) const {
    const float4* b = &q[i * QB];
    if constexpr (COMPACT) {
      const int r1 = r & 3, r2 = r >> 2;
      const float2 lo = r1 == 1 ? make_float2(b[0].x, b[0].y)
                        : r1 == 2 ? make_float2(b[0].z, b[0].w)
                                  : make_float2(b[1].x, b[1].y);
      float2 hi = make_float2(b[1].z, b[1].w);  // r2 == 1
      if constexpr (CB == 3) {
        if (r2 == 2) hi = make_float2(b[2].x, b[2].y);
        if (r2 == 3) hi = make_float2(b[2].z, b[2].w);
      }
      if (r2 == 0) return lo;
      if (r1 == 0) return hi;
      return cmul(lo, hi);
    } else if constexpr (SHIFT) {
      const float4 p = b[(r - 1) / 2];
      return ((r - 1) & 1) ? make_float2(p.z, p.w) : make_float2(p.x, p.y);
    } else {
      const float4 p = b[r / 2];
      return (r & 1) ? make_float2(p.z, p.w) : make_float2(p.x, p.y);
    }
  }
};

template <int L, bool REV, int s, int CM>
__device__ __forceinline__ void sfft_load_tw(StageTw<L, REV, s, CM>& tw, int t,
                                             const float2* __restrict__ stw) {
  using P = PlanOf<L, REV>;
  using W = StageTw<L, REV, s, CM>;
  if constexpr (W::Ns > 1) {
    constexpr int M = P::M(s);
#pragma unroll
    for (int i = 0; i < W::NB; ++i) {
      const int b = t + i * P::T;
      if ((M % P::T) != 0 && b >= M) continue;
      const float2* row = stw + P::TWOFF(s) + (b % W::Ns) * W::RP;
      if constexpr (W::COMPACT) {
        // rows are an even number of float2 apart, so entries 2-3 form an
        // aligned float4
        const float2 e1 = row[1], e4 = row[4];
        const float4 e23 = *reinterpret_cast<const float4*>(row + 2);
        tw.q[i * W::QB + 0] = make_float4(e1.x, e1.y, e23.x, e23.y);
        tw.q[i * W::QB + 1] = make_float4(e23.z, e23.w, e4.x, e4.y);
        if constexpr (W::CB == 3) {
          const float2 e8 = row[8];
          const float2 e12 = W::R > 12 ? row[12] : make_float2(1.f, 0.f);
          tw.q[i * W::QB + 2] = make_float4(e8.x, e8.y, e12.x, e12.y);
        }
      } else if constexpr (W::SHIFT) {
#pragma unroll
        for (int m = 0; m < W::QB; ++m) {
          const float2 ea = row[2 * m + 1], eb = row[2 * m + 2];
          tw.q[i * W::QB + m] = make_float4(ea.x, ea.y, eb.x, eb.y);
        }
      } else {
        const float4* q = reinterpret_cast<const float4*>(row);
#pragma unroll
        for (int r2 = 0; r2 < W::RP / 2; ++r2) tw.q[i * (W::RP / 2) + r2] = q[r2];
      }
    }
  }
}

// The twiddle multiply of one stage (lets two transforms share one twiddle
// set that dies before either butterfly runs).
template <int L, bool REV, int DIR, int s, int MAXE, int CM>
__device__ __forceinline__ void sfft_twiddle(float2 (&v)[MAXE], int t,
                                             const StageTw<L, REV, s, CM>& tw) {
  using P = PlanOf<L, REV>;
  using W = StageTw<L, REV, s, CM>;
  constexpr int R = P::R(s), M = P::M(s), NB = P::NB(s);
  if constexpr (W::Ns > 1) {
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int b = t + i * P::T;
      if ((M % P::T) != 0 && b >= M) continue;
      float2* x = &v[i * R];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const float2 w = tw.get(i, r);
        x[r] = DIR < 0 ? cmul(x[r], w) : cmulc(x[r], w);
      }
    }
  }
}

// sfft_compute with the twiddles already in registers
template <int L, bool REV, int DIR, int s, int MAXE, int CM>
__device__ __forceinline__ void sfft_compute_tw(float2 (&v)[MAXE], int t,
                                                const StageTw<L, REV, s, CM>& tw) {
  using P = PlanOf<L, REV>;
  constexpr int R = P::R(s), M = P::M(s), NB = P::NB(s);
  sfft_twiddle<L, REV, DIR, s>(v, t, tw);
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = t + i * P::T;
    if ((M % P::T) != 0 && b >= M) continue;
    Dft<R, DIR>::run(&v[i * R]);
  }
}

template <int L, bool REV, int DIR, int s, int MAXE>
__device__ __forceinline__ void sfft_butterflies(float2 (&v)[MAXE], int t) {
  using P = PlanOf<L, REV>;
  constexpr int R = P::R(s), M = P::M(s), NB = P::NB(s);
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = t + i * P::T;
    if ((M % P::T) != 0 && b >= M) continue;
    Dft<R, DIR>::run(&v[i * R]);
  }
}

// Stages 1..S-1 (stage 0 inputs must already be in v); the LDS region is
// reused, so the caller must not have other threads reading it.  A stage's
// twiddles are requested before the barriers of its exchange (the barrier
// waits for LDS traffic only), so their latency overlaps the exchange; the
// data registers are free at that point.
template <int L, bool REV, int DIR, int s, int CM, int MAXE, class Addr>
__device__ __forceinline__ void sfft_rest(float2 (&v)[MAXE], int t, float2* lds, Addr addr,
                                          const float2* __restrict__ tw) {
  using P = PlanOf<L, REV>;
  if constexpr (s < P::S) {
    sfft_store_lds<L, REV, s - 1>(v, t, lds, addr);
    StageTw<L, REV, s, CM> stw;
    sfft_load_tw<L, REV, s>(stw, t, tw);
    __syncthreads();
    sfft_load_lds<L, REV, s>(v, t, lds, addr);
    __syncthreads();
    sfft_compute_tw<L, REV, DIR, s>(v, t, stw);
    sfft_rest<L, REV, DIR, s + 1, CM>(v, t, lds, addr, tw);
  }
}

// Whole transform: stage 0 compute, then the remaining stages.  CM is the
// smallest radix whose stages keep compact twiddles (StageTw); the x-R2C
// passes measured faster with full rows at radix 8 (512 x 512 x 256: 105
// against 110 us) and with compact ones at radix 16 (256 x 256 x 128).
template <int L, bool REV, int DIR, int CM = FC_TW_COMPACT_MIN, int MAXE, class Addr>
__device__ __forceinline__ void sfft_run(float2 (&v)[MAXE], int t, float2* lds, Addr addr,
                                         const float2* __restrict__ tw) {
  sfft_compute<L, REV, DIR, 0>(v, t, tw);
  sfft_rest<L, REV, DIR, 1, CM>(v, t, lds, addr, tw);
}

}  // namespace fcamd
