// plane_kernels.hpp -- x and y transforms of a whole z-plane in one
// workgroup, for volumes whose half-spectrum plane fits in LDS.
//
// The plain path runs x-R2C and y-FFT as two launches, each one HBM round
// trip of the spectrum; for small volumes (128 x 128 x 64: 8-10 MB per
// pass) each of those launches costs 5-9 us of ramp, latency and drain
// rather than bandwidth.  Here one workgroup per z-plane reads the plane's
// real rows once, transforms its rows (two real rows per complex FFT, as the
// x passes do) into a half-spectrum plane held in LDS, transforms that
// plane's columns in place and writes the spectrum plane: the forward x and
// y passes in one launch.  The inverse does y-IFFT and x-C2R the same way.
// Workgroups past the image's planes transform the kernel's planes (taps
// placed circularly, x at (x - cx) mod nx, y at (y - cy) mod ny, as
// fftShiftKernel places them, src/convolution3Dfft.cu:128-166) into the
// kernel-plane layout the fused z pass reads.
//
// LDS: the plane P[ny][PB] (PB = nx/2 + 2 float2) and the row-pair exchange
// rows.  Real rows are staged in P with a pitch of nx + 4 floats = 2 * PB, so
// real row y lies inside spectrum row y: a row pair only ever touches its
// own two rows of P.  Threads without a transform in a phase exchange
// through one dummy LDS slot, so every thread reaches every barrier.
#pragma once

#include "static_kernels_impl.hpp"

namespace fcamd {

template <int NX, int NY>
struct PlaneGeom {
  static constexpr int TX = threads_per_transform<NX>();  // per row pair
  static constexpr int TY = threads_per_transform<NY>();  // per column
  static constexpr int NB = NX / 2 + 1;                   // half-spectrum bins
  static constexpr int PB = NB + 1;                       // plane row pitch (float2)
  static constexpr int SL = 2 * PB;                       // real staging pitch (floats)
  static constexpr int PAIRS = NY / 2;
  static constexpr int LPX = XRow<NX>::LP;
  static constexpr int want(int a, int b) { return a > b ? a : b; }
  static constexpr int least(int a, int b) { return a < b ? a : b; }
  static constexpr int NT_ = (want(TX * PAIRS, TY * NB) + 63) / 64 * 64;
  static constexpr int NT = NT_ > 1024 ? 1024 : NT_;
  static constexpr int P_FLOAT2 = NY * PB + 1;  // + dummy slot
  // row pairs at once: as many as the threads and the LDS left by P allow
  static constexpr int XC_LDS = (160 * 1024 / 8 - P_FLOAT2 - 1) / LPX;
  static constexpr int XC = least(least(NT / TX, PAIRS), XC_LDS);
  static constexpr int YC = NT / TY < NB ? NT / TY : NB;  // columns at once
  static constexpr int R_FLOAT2 = XC * LPX + 1;           // + dummy slot
  static constexpr long LDS_BYTES = 8L * (P_FLOAT2 + R_FLOAT2);
  static_assert(NX % 4 == 0 && NY % 2 == 0, "plane kernels take nx % 4 == 0, even ny");
  static_assert(XC >= 1 && YC >= 1, "plane too wide for one workgroup");
  static_assert(LDS_BYTES <= 160 * 1024, "plane does not fit in LDS");
};

// x phase forward: row pairs of P's staged real rows (image) or of the
// kernel's taps -> half spectra in P.  `image` is workgroup-uniform.
template <int NX, int NY>
__device__ __forceinline__ void plane_x_forward(const PlaneArgs& a, bool image, int plane,
                                                float2* P, float2* rows) {
  using G = PlaneGeom<NX, NY>;
  using PX = PlanOf<NX, false>;
  const float* st = reinterpret_cast<const float*>(P);
  const int c = threadIdx.x / G::TX, t0 = threadIdx.x % G::TX;
  const bool lane = c < G::XC;
  const int t = lane ? t0 : 0;
  for (int g0 = 0; g0 < G::PAIRS; g0 += G::XC) {
    const int pr = g0 + c;
    const bool act = lane && pr < G::PAIRS;
    const int ya = 2 * pr, yb = ya + 1;
    float2 v[PX::MAXE()];
    if (image) {
      sfft_for_inputs<NX, false>(t, [&](int slot, int pos) {
        v[slot] = act ? make_float2(st[ya * G::SL + pos], st[yb * G::SL + pos]) : make_float2(0.f, 0.f);
      });
    } else {
      // row y of kernel plane `plane` holds kernel row (y + cy) mod NY
      int ia = ya + a.cy, ib = yb + a.cy;
      if (ia >= NY) ia -= NY;
      if (ib >= NY) ib -= NY;
      const bool oka = act && ia < a.ky, okb = act && ib < a.ky;
      const float* ka = a.kernel + ((long)plane * a.ky + (oka ? ia : 0)) * a.kx;
      const float* kb = a.kernel + ((long)plane * a.ky + (okb ? ib : 0)) * a.kx;
      const int cx = a.cx, kx = a.kx;
      sfft_for_inputs<NX, false>(t, [&](int slot, int pos) {
        int x = pos + cx;
        if (x >= NX) x -= NX;
        float xa = 0.f, xb = 0.f;
        if (x < kx) {
          if (oka) xa = ka[x];
          if (okb) xb = kb[x];
        }
        v[slot] = make_float2(xa, xb);
      });
    }
    auto addr = [&](int pos) { return act ? c * G::LPX + XRow<NX>::at(pos) : G::R_FLOAT2 - 1; };
    sfft_run<NX, false, -1>(v, t, rows, addr, a.fx.stw);
    sfft_for_outputs<NX, false>(t, [&](int slot, int pos) { rows[addr(pos)] = v[slot]; });
    __syncthreads();
    if (act) {
      const float2* row = rows + c * G::LPX;
      for (int kb = t; kb < G::NB; kb += G::TX) {
        const float2 Z = row[XRow<NX>::at(kb)];
        const float2 Zm = row[XRow<NX>::at(kb == 0 ? 0 : NX - kb)];
        P[ya * G::PB + kb] = make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
        P[yb * G::PB + kb] = make_float2(0.5f * (Z.y + Zm.y), 0.5f * (Zm.x - Z.x));
      }
    }
    __syncthreads();  // the exchange rows are reused by the next group
  }
}

// y phase: FFT (DIR) of every column of P in place; the results go to
// `out` (a spectrum plane of pitch nxp) when non-null, else back into P.
template <int NX, int NY, int DIR>
__device__ __forceinline__ void plane_y(const PlaneArgs& a, float2* P, float2* out) {
  using G = PlaneGeom<NX, NY>;
  using PY = PlanOf<NY, false>;
  const int c = threadIdx.x % G::YC, t0 = threadIdx.x / G::YC;
  const bool lane = t0 < G::TY;
  const int t = lane ? t0 : 0;
  const long nxp = a.nxp;
  for (int g0 = 0; g0 < G::NB; g0 += G::YC) {
    const int col = g0 + c;
    const bool act = lane && col < G::NB;
    auto addr = [&](int pos) { return act ? pos * G::PB + col : G::P_FLOAT2 - 1; };
    float2 v[PY::MAXE()];
    sfft_for_inputs<NY, false>(t, [&](int slot, int pos) { v[slot] = P[addr(pos)]; });
    __syncthreads();  // every input of the column is read before the exchange
    sfft_run<NY, false, DIR>(v, t, P, addr, a.fy.stw);
    if (out) {
      if (act) sfft_for_outputs<NY, false>(t, [&](int slot, int pos) { out[(long)pos * nxp + col] = v[slot]; });
    } else {
      sfft_for_outputs<NY, false>(t, [&](int slot, int pos) { P[addr(pos)] = v[slot]; });
    }
  }
}

// Forward: workgroup z < nz transforms image plane z into spec; the next kz
// workgroups transform the kernel's planes into kplanes.
template <int NX, int NY>
__global__ void __launch_bounds__((PlaneGeom<NX, NY>::NT)) sk_plane_fwd(PlaneArgs a) {
  using G = PlaneGeom<NX, NY>;
  __shared__ float2 P[G::P_FLOAT2];
  __shared__ float2 rows[G::R_FLOAT2];
  const bool image = (int)blockIdx.x < a.nz;
  const int plane = image ? (int)blockIdx.x : (int)blockIdx.x - a.nz;
  if (image) {
    // the plane's real rows, staged with a pitch of SL floats
    float* st = reinterpret_cast<float*>(P);
    const float* g = a.src + (long)plane * NX * NY;
    if ((reinterpret_cast<uintptr_t>(g) & 15) == 0) {
      constexpr int Q = NX / 4, N4 = Q * NY, IT = (N4 + G::NT - 1) / G::NT;
      const float4* g4 = reinterpret_cast<const float4*>(g);
      float4 r[IT];
#pragma unroll
      for (int j = 0; j < IT; ++j) {
        const int i = threadIdx.x + j * G::NT;
        r[j] = i < N4 ? g4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
      }
#pragma unroll
      for (int j = 0; j < IT; ++j) {
        const int i = threadIdx.x + j * G::NT;
        if (i < N4) {
          const int y = i / Q;
          *reinterpret_cast<float4*>(st + y * G::SL + 4 * (i - y * Q)) = r[j];
        }
      }
    } else {
      for (int i = threadIdx.x; i < NX * NY; i += G::NT) {
        const int y = i / NX;
        st[y * G::SL + (i - y * NX)] = g[i];
      }
    }
    __syncthreads();
  }
  plane_x_forward<NX, NY>(a, image, plane, P, rows);
  float2* out = (image ? a.spec : a.kplanes) + (long)plane * NY * a.nxp;
  plane_y<NX, NY, -1>(a, P, out);
}

// Inverse: workgroup z transforms spectrum plane z into real plane z.
template <int NX, int NY>
__global__ void __launch_bounds__((PlaneGeom<NX, NY>::NT)) sk_plane_inv(PlaneArgs a) {
  using G = PlaneGeom<NX, NY>;
  using PX = PlanOf<NX, false>;
  __shared__ float2 P[G::P_FLOAT2];
  __shared__ float2 rows[G::R_FLOAT2];
  const int plane = blockIdx.x;
  {
    const float2* sp = a.spec + (long)plane * NY * a.nxp;
    constexpr int N = NY * G::NB, IT = (N + G::NT - 1) / G::NT;
    float2 r[IT];
#pragma unroll
    for (int j = 0; j < IT; ++j) {
      const int i = threadIdx.x + j * G::NT;
      const int y = i / G::NB;
      r[j] = i < N ? sp[(long)y * a.nxp + (i - y * G::NB)] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int j = 0; j < IT; ++j) {
      const int i = threadIdx.x + j * G::NT;
      const int y = i / G::NB;
      if (i < N) P[y * G::PB + (i - y * G::NB)] = r[j];
    }
  }
  __syncthreads();
  plane_y<NX, NY, +1>(a, P, nullptr);
  __syncthreads();
  // x-C2R of row pairs; the real rows go to the staging pitch inside their
  // own spectrum rows
  float* st = reinterpret_cast<float*>(P);
  const int c = threadIdx.x / G::TX, t0 = threadIdx.x % G::TX;
  const bool lane = c < G::XC;
  const int t = lane ? t0 : 0;
  for (int g0 = 0; g0 < G::PAIRS; g0 += G::XC) {
    const int pr = g0 + c;
    const bool act = lane && pr < G::PAIRS;
    const int ya = 2 * pr, yb = ya + 1;
    float2 v[PX::MAXE()];
    sfft_for_inputs<NX, false>(t, [&](int slot, int pos) {
      const bool mirror = pos >= G::NB;
      const int kb = mirror ? NX - pos : pos;
      float2 A = act ? P[ya * G::PB + kb] : make_float2(0.f, 0.f);
      float2 B = act ? P[yb * G::PB + kb] : make_float2(0.f, 0.f);
      if (kb == 0 || 2 * kb == NX) {
        A.y = 0.f;
        B.y = 0.f;
      }
      if (mirror) {
        A.y = -A.y;
        B.y = -B.y;
      }
      v[slot] = make_float2(A.x - B.y, A.y + B.x);
    });
    auto addr = [&](int pos) { return act ? c * G::LPX + XRow<NX>::at(pos) : G::R_FLOAT2 - 1; };
    sfft_run<NX, false, +1>(v, t, rows, addr, a.fx.stw);
    if (act)
      sfft_for_outputs<NX, false>(t, [&](int slot, int pos) {
        st[ya * G::SL + pos] = v[slot].x;
        st[yb * G::SL + pos] = v[slot].y;
      });
    __syncthreads();  // the exchange rows are reused by the next group
  }
  float* g = a.dst + (long)plane * NX * NY;
  if ((reinterpret_cast<uintptr_t>(g) & 15) == 0) {
    constexpr int Q = NX / 4, N4 = Q * NY, IT = (N4 + G::NT - 1) / G::NT;
    float4* g4 = reinterpret_cast<float4*>(g);
#pragma unroll
    for (int j = 0; j < IT; ++j) {
      const int i = threadIdx.x + j * G::NT;
      if (i < N4) {
        const int y = i / Q;
        g4[i] = *reinterpret_cast<const float4*>(st + y * G::SL + 4 * (i - y * Q));
      }
    }
  } else {
    for (int i = threadIdx.x; i < NX * NY; i += G::NT) {
      const int y = i / NX;
      g[i] = st[y * G::SL + (i - y * NX)];
    }
  }
}

// ---------------------------------------------------------------------------
// Split kernels: two workgroups per plane that share no work.
//
// One workgroup per plane puts a plane's x and y transforms on one CU, where
// they are VALU-bound (about 0.6 M lane operations for a 128 x 128 plane,
// ~8 us), and with fewer planes than CUs the other CUs idle.  Here each plane
// takes two workgroups, each with half the transforms' work:
//  - forward, part p computes the x bins k = p (mod 2).  A radix-2
//    decimation in frequency turns the row pair's length-nx FFT into one of
//    length h = nx/2 over u_p[n] = (z[n] + (-1)^p z[n + h]) W_nx^(p n),
//    whose output m is bin 2m + p.  The two-real-rows split of bin k needs
//    bin nx - k, of the same parity (nx % 4 == 0), so it stays in the part.
//    Part p then transforms its columns and writes spectrum bins 2j + p.
//  - inverse, part q computes the rows y = q (mod 2): outputs 2m + q of the
//    length-ny column IFFT are the length-g = ny/2 IFFT of
//    (X[k] + (-1)^q X[k + g]) W_ny^(-q k), formed as the plane is read; the
//    x-C2R then runs on the part's g rows (ny % 4 == 0 keeps them in pairs).
// The parts of a plane take blockIdx values equal mod 8, which the hardware
// deals to one XCD, so the plane read twice and the interleaved bins written
// by the two parts meet in that XCD's L2.
// ---------------------------------------------------------------------------
template <int S>
__device__ __forceinline__ void plane_decode(int* q, int* part) {
  const int b = blockIdx.x, r = b / 8;
  *part = r % S;
  *q = (r / S) * 8 + b % 8;
}

template <int S>
inline int plane_blocks(int nplanes) {
  return (nplanes + 7) / 8 * 8 * S;
}

template <int NX, int NY>
struct PlaneGeomF2 {
  static constexpr int H = NX / 2;                       // row FFT length
  static constexpr int TX = threads_per_transform<H>();  // per row pair
  static constexpr int TY = threads_per_transform<NY>();
  static constexpr int NC = H / 2 + 1;  // columns of part 0 (part 1: H / 2)
  static constexpr int PB = H + 2;      // plane row pitch (float2)
  static constexpr int SL = 2 * PB;     // real staging pitch (floats)
  static constexpr int PAIRS = NY / 2;
  static constexpr int LPX = XRow<H>::LP;
  static constexpr int want(int a, int b) { return a > b ? a : b; }
  static constexpr int least(int a, int b) { return a < b ? a : b; }
  static constexpr int NT_ = (want(TX * PAIRS, TY * NC) + 63) / 64 * 64;
  static constexpr int NT = NT_ > 1024 ? 1024 : NT_;
  static constexpr int P_FLOAT2 = NY * PB + 1;  // + dummy slot
  static constexpr int XC_LDS = (160 * 1024 / 8 - P_FLOAT2 - 1) / LPX;
  static constexpr int XC = least(least(NT / TX, PAIRS), XC_LDS);
  static constexpr int YC = least(NT / TY, NC);
  static constexpr int R_FLOAT2 = XC * LPX + 1;  // + dummy slot
  static constexpr long LDS_BYTES = 8L * (P_FLOAT2 + R_FLOAT2);
  static_assert(NX % 4 == 0 && NY % 2 == 0, "split forward takes nx % 4 == 0, even ny");
  static_assert(XC >= 1 && YC >= 1, "plane too wide for one workgroup");
  static_assert(LDS_BYTES <= 160 * 1024, "plane does not fit in LDS");
};

template <int NX, int NY>
__global__ void __launch_bounds__((PlaneGeomF2<NX, NY>::NT)) sk_plane_fwd2(PlaneArgs a) {
  using G = PlaneGeomF2<NX, NY>;
  using PX = PlanOf<G::H, false>;
  using PY = PlanOf<NY, false>;
  __shared__ float2 P[G::P_FLOAT2];
  __shared__ float2 rows[G::R_FLOAT2];
  int q, p;
  plane_decode<2>(&q, &p);
  if (q >= a.nz + (a.kernel ? a.kz : 0)) return;  // workgroup-uniform, before any barrier
  const bool image = q < a.nz;
  const int plane = image ? q : q - a.nz;
  const int c = threadIdx.x / G::TX, t0 = threadIdx.x % G::TX;
  const bool lane = c < G::XC;
  const int t = lane ? t0 : 0;
  // W_nx^n of the thread's inputs n (part 1), requested before the plane
  float2 w[PX::MAXE()];
  sfft_for_inputs<G::H, false>(t, [&](int slot, int pos) { w[slot] = p ? a.fx.tw[pos] : make_float2(1.f, 0.f); });
  float* st = reinterpret_cast<float*>(P);
  if (image) {
    const float* g = a.src + (long)plane * NX * NY;
    if ((reinterpret_cast<uintptr_t>(g) & 15) == 0) {
      constexpr int Q = NX / 4, N4 = Q * NY, IT = (N4 + G::NT - 1) / G::NT;
      const float4* g4 = reinterpret_cast<const float4*>(g);
      float4 r[IT];
#pragma unroll
      for (int j = 0; j < IT; ++j) {
        const int i = threadIdx.x + j * G::NT;
        r[j] = i < N4 ? g4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
      }
#pragma unroll
      for (int j = 0; j < IT; ++j) {
        const int i = threadIdx.x + j * G::NT;
        if (i < N4) {
          const int y = i / Q;
          *reinterpret_cast<float4*>(st + y * G::SL + 4 * (i - y * Q)) = r[j];
        }
      }
    } else {
      for (int i = threadIdx.x; i < NX * NY; i += G::NT) {
        const int y = i / NX;
        st[y * G::SL + (i - y * NX)] = g[i];
      }
    }
    __syncthreads();
  }
  const int ncols = G::NC - p;
  for (int g0 = 0; g0 < G::PAIRS; g0 += G::XC) {
    const int pr = g0 + c;
    const bool act = lane && pr < G::PAIRS;
    const int ya = 2 * pr, yb = ya + 1;
    float2 v[PX::MAXE()];
    if (image) {
      sfft_for_inputs<G::H, false>(t, [&](int slot, int pos) {
        float2 z0 = make_float2(0.f, 0.f), z1 = z0;
        if (act) {
          z0 = make_float2(st[ya * G::SL + pos], st[yb * G::SL + pos]);
          z1 = make_float2(st[ya * G::SL + pos + G::H], st[yb * G::SL + pos + G::H]);
        }
        v[slot] = p ? cmul(csub(z0, z1), w[slot]) : cadd(z0, z1);
      });
    } else {
      // row y of kernel plane `plane` holds kernel row (y + cy) mod NY, and
      // its position x tap (x + cx) mod NX
      int ia = ya + a.cy, ib = yb + a.cy;
      if (ia >= NY) ia -= NY;
      if (ib >= NY) ib -= NY;
      const bool oka = act && ia < a.ky, okb = act && ib < a.ky;
      const float* ka = a.kernel + ((long)plane * a.ky + (oka ? ia : 0)) * a.kx;
      const float* kb = a.kernel + ((long)plane * a.ky + (okb ? ib : 0)) * a.kx;
      const int cx = a.cx, kx = a.kx;
      auto tap = [&](int pos) {
        int x = pos + cx;
        if (x >= NX) x -= NX;
        float xa = 0.f, xb = 0.f;
        if (x < kx) {
          if (oka) xa = ka[x];
          if (okb) xb = kb[x];
        }
        return make_float2(xa, xb);
      };
      sfft_for_inputs<G::H, false>(t, [&](int slot, int pos) {
        const float2 z0 = tap(pos), z1 = tap(pos + G::H);
        v[slot] = p ? cmul(csub(z0, z1), w[slot]) : cadd(z0, z1);
      });
    }
    auto addr = [&](int pos) { return act ? c * G::LPX + XRow<G::H>::at(pos) : G::R_FLOAT2 - 1; };
    sfft_run<G::H, false, -1>(v, t, rows, addr, a.hx_stw);
    sfft_for_outputs<G::H, false>(t, [&](int slot, int pos) { rows[addr(pos)] = v[slot]; });
    __syncthreads();
    if (act) {
      // local column j is bin k = 2j + p; bin nx - k is output jm
      const float2* row = rows + c * G::LPX;
      for (int j = t; j < ncols; j += G::TX) {
        const int jm = p ? G::H - 1 - j : (j == 0 ? 0 : G::H - j);
        const float2 Z = row[XRow<G::H>::at(j)];
        const float2 Zm = row[XRow<G::H>::at(jm)];
        P[ya * G::PB + j] = make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
        P[yb * G::PB + j] = make_float2(0.5f * (Z.y + Zm.y), 0.5f * (Zm.x - Z.x));
      }
    }
    __syncthreads();  // the exchange rows are reused by the next group
  }
  // columns of the part: FFT along y, bins 2j + p of the spectrum plane
  float2* out = (image ? a.spec : a.kplanes) + (long)plane * NY * a.nxp + p;
  const long nxp = a.nxp;
  const int cc = threadIdx.x % G::YC, u0 = threadIdx.x / G::YC;
  const bool ylane = u0 < G::TY;
  const int u = ylane ? u0 : 0;
  for (int g0 = 0; g0 < ncols; g0 += G::YC) {
    const int col = g0 + cc;
    const bool act = ylane && col < ncols;
    auto addr = [&](int pos) { return act ? pos * G::PB + col : G::P_FLOAT2 - 1; };
    float2 v[PY::MAXE()];
    sfft_for_inputs<NY, false>(u, [&](int slot, int pos) { v[slot] = P[addr(pos)]; });
    __syncthreads();  // every input of the column is read before the exchange
    sfft_run<NY, false, -1>(v, u, P, addr, a.fy.stw);
    if (act)
      sfft_for_outputs<NY, false>(u, [&](int slot, int pos) { out[(long)pos * nxp + 2 * col] = v[slot]; });
  }
}

template <int NX, int NY>
struct PlaneGeomI2 {
  static constexpr int GY = NY / 2;                       // column IFFT length
  static constexpr int TX = threads_per_transform<NX>();  // per row pair
  static constexpr int TY = threads_per_transform<GY>();
  static constexpr int NB = NX / 2 + 1;
  static constexpr int PB = NB + 1;
  static constexpr int SL = 2 * PB;
  static constexpr int PAIRS = GY / 2;
  static constexpr int LPX = XRow<NX>::LP;
  static constexpr int want(int a, int b) { return a > b ? a : b; }
  static constexpr int least(int a, int b) { return a < b ? a : b; }
  static constexpr int NT_ = (want(TX * PAIRS, TY * NB) + 63) / 64 * 64;
  static constexpr int NT = NT_ > 1024 ? 1024 : NT_;
  static constexpr int P_FLOAT2 = GY * PB + 1;  // + dummy slot
  static constexpr int XC_LDS = (160 * 1024 / 8 - P_FLOAT2 - 1) / LPX;
  static constexpr int XC = least(least(NT / TX, PAIRS), XC_LDS);
  static constexpr int YC = least(NT / TY, NB);
  static constexpr int R_FLOAT2 = XC * LPX + 1;  // + dummy slot
  static constexpr long LDS_BYTES = 8L * (P_FLOAT2 + R_FLOAT2);
  static_assert(NX % 4 == 0 && NY % 4 == 0, "split inverse takes nx % 4 == 0, ny % 4 == 0");
  static_assert(XC >= 1 && YC >= 1, "plane too wide for one workgroup");
  static_assert(LDS_BYTES <= 160 * 1024, "plane does not fit in LDS");
};

template <int NX, int NY>
__global__ void __launch_bounds__((PlaneGeomI2<NX, NY>::NT)) sk_plane_inv2(PlaneArgs a) {
  using G = PlaneGeomI2<NX, NY>;
  using PX = PlanOf<NX, false>;
  using PY = PlanOf<G::GY, false>;
  __shared__ float2 P[G::P_FLOAT2];
  __shared__ float2 rows[G::R_FLOAT2];
  int plane, q;
  plane_decode<2>(&plane, &q);
  if (plane >= a.nz) return;  // workgroup-uniform, before any barrier
  {
    // row k of P: X[k] + X[k + g] (part 0) or (X[k] - X[k + g]) W_ny^-k
    const float2* sp = a.spec + (long)plane * NY * a.nxp;
    const long half = (long)G::GY * a.nxp;
    constexpr int N = G::GY * G::NB, IT = (N + G::NT - 1) / G::NT;
    float2 r0[IT], r1[IT], w[IT];
#pragma unroll
    for (int j = 0; j < IT; ++j) {
      const int i = threadIdx.x + j * G::NT;
      const int k = i / G::NB;
      const long o = (long)k * a.nxp + (i - k * G::NB);
      r0[j] = i < N ? sp[o] : make_float2(0.f, 0.f);
      r1[j] = i < N ? sp[o + half] : make_float2(0.f, 0.f);
      w[j] = q && i < N ? a.fy.tw[k] : make_float2(1.f, 0.f);
    }
#pragma unroll
    for (int j = 0; j < IT; ++j) {
      const int i = threadIdx.x + j * G::NT;
      const int k = i / G::NB;
      if (i < N) P[k * G::PB + (i - k * G::NB)] = q ? cmulc(csub(r0[j], r1[j]), w[j]) : cadd(r0[j], r1[j]);
    }
  }
  __syncthreads();
  {
    // the columns' length-g IFFTs, in place
    const int cc = threadIdx.x % G::YC, u0 = threadIdx.x / G::YC;
    const bool ylane = u0 < G::TY;
    const int u = ylane ? u0 : 0;
    for (int g0 = 0; g0 < G::NB; g0 += G::YC) {
      const int col = g0 + cc;
      const bool act = ylane && col < G::NB;
      auto addr = [&](int pos) { return act ? pos * G::PB + col : G::P_FLOAT2 - 1; };
      float2 v[PY::MAXE()];
      sfft_for_inputs<G::GY, false>(u, [&](int slot, int pos) { v[slot] = P[addr(pos)]; });
      __syncthreads();  // every input of the column is read before the exchange
      sfft_run<G::GY, false, +1>(v, u, P, addr, a.hy_stw);
      sfft_for_outputs<G::GY, false>(u, [&](int slot, int pos) { P[addr(pos)] = v[slot]; });
    }
  }
  __syncthreads();
  // x-C2R of the part's row pairs; the real rows go to the staging pitch
  // inside their own spectrum rows
  float* st = reinterpret_cast<float*>(P);
  const int c = threadIdx.x / G::TX, t0 = threadIdx.x % G::TX;
  const bool lane = c < G::XC;
  const int t = lane ? t0 : 0;
  for (int g0 = 0; g0 < G::PAIRS; g0 += G::XC) {
    const int pr = g0 + c;
    const bool act = lane && pr < G::PAIRS;
    const int ya = 2 * pr, yb = ya + 1;
    float2 v[PX::MAXE()];
    sfft_for_inputs<NX, false>(t, [&](int slot, int pos) {
      const bool mirror = pos >= G::NB;
      const int kb = mirror ? NX - pos : pos;
      float2 A = act ? P[ya * G::PB + kb] : make_float2(0.f, 0.f);
      float2 B = act ? P[yb * G::PB + kb] : make_float2(0.f, 0.f);
      if (kb == 0 || 2 * kb == NX) {
        A.y = 0.f;
        B.y = 0.f;
      }
      if (mirror) {
        A.y = -A.y;
        B.y = -B.y;
      }
      v[slot] = make_float2(A.x - B.y, A.y + B.x);
    });
    auto addr = [&](int pos) { return act ? c * G::LPX + XRow<NX>::at(pos) : G::R_FLOAT2 - 1; };
    sfft_run<NX, false, +1>(v, t, rows, addr, a.fx.stw);
    if (act)
      sfft_for_outputs<NX, false>(t, [&](int slot, int pos) {
        st[ya * G::SL + pos] = v[slot].x;
        st[yb * G::SL + pos] = v[slot].y;
      });
    __syncthreads();  // the exchange rows are reused by the next group
  }
  // local row m is image row 2m + q
  float* g = a.dst + (long)plane * NX * NY + (long)q * NX;
  if ((reinterpret_cast<uintptr_t>(g) & 15) == 0) {
    constexpr int Q = NX / 4, N4 = Q * G::GY, IT = (N4 + G::NT - 1) / G::NT;
#pragma unroll
    for (int j = 0; j < IT; ++j) {
      const int i = threadIdx.x + j * G::NT;
      if (i < N4) {
        const int m = i / Q, x4 = i - m * Q;
        *reinterpret_cast<float4*>(g + (long)2 * m * NX + 4 * x4) =
            *reinterpret_cast<const float4*>(st + m * G::SL + 4 * x4);
      }
    }
  } else {
    for (int i = threadIdx.x; i < NX * G::GY; i += G::NT) {
      const int m = i / NX, x = i - m * NX;
      g[(long)2 * m * NX + x] = st[m * G::SL + x];
    }
  }
}

// ---------------------------------------------------------------------------
// Wide split kernels: planes whose half spectrum does not fit one
// workgroup's LDS (256 x 256: 264 KB) but whose even or odd half does.  The
// decomposition is the split kernels'; three changes keep a part within
// 160 KB of LDS and a single workgroup:
//  - the real rows are not staged: a row pair's length-nx/2 input u_p is
//    formed from coalesced global loads of its two rows;
//  - a row pair's FFT exchanges in place in the pair's own two rows of P
//    (rows ya and yb are adjacent, 2 * PB float2 >= the padded exchange),
//    with a barrier between reading the inputs and the first exchange;
//  - two columns of a part that are real share one complex transform, so
//    a part's columns take nx/4 transforms: forward bins 0 and nx/2 of part
//    0 (exactly real after the two-real-rows separation), separated after
//    the FFT by the same even/odd split; inverse columns 0 and nx/2, whose
//    y-IFFT outputs are real for the Hermitian spectrum of a real volume:
//    the real part of the packed transform is column 0's, the imaginary
//    part column nx/2's (the x-C2R takes only the real part of those two
//    columns anyway).
// ---------------------------------------------------------------------------
template <int NX, int NY>
struct PlaneGeomF2W {
  static constexpr int want(int a, int b) { return a > b ? a : b; }
  static constexpr int least(int a, int b) { return a < b ? a : b; }
  static constexpr int H = NX / 2;                       // row FFT length
  static constexpr int TX = threads_per_transform<H>();  // per row pair
  static constexpr int TY = threads_per_transform<NY>();
  static constexpr int NC = H / 2 + 1;  // columns of part 0 (part 1: H / 2)
  static constexpr int NCT = NC - 1;    // column transforms of either part
  static constexpr int LPH = XRow<H>::LP;
  static constexpr int PB = want(NC, (LPH + 1) / 2);  // two rows hold a pair's exchange
  static constexpr int PAIRS = NY / 2;
  static constexpr int NT_ = (want(TX * PAIRS, TY * NCT) + 63) / 64 * 64;
  static constexpr int NT = NT_ > 1024 ? 1024 : NT_;
  static constexpr int XC = least(NT / TX, PAIRS);
  static constexpr int YC = least(NT / TY, NCT);
  static constexpr int JX = (NC + TX - 1) / TX;  // separated bins per thread
  static constexpr int P_FLOAT2 = NY * PB + 1;   // + dummy slot
  static constexpr long LDS_BYTES = 8L * P_FLOAT2;
  static_assert(NX % 8 == 0 && NY % 2 == 0, "wide split forward takes nx % 8 == 0, even ny");
  static_assert(XC >= 1 && YC >= 1, "plane too wide for one workgroup");
  static_assert(LDS_BYTES <= 160 * 1024, "half plane does not fit in LDS");
};

template <int NX, int NY>
__global__ void __launch_bounds__((PlaneGeomF2W<NX, NY>::NT)) sk_plane_fwd2w(PlaneArgs a) {
  using G = PlaneGeomF2W<NX, NY>;
  using PX = PlanOf<G::H, false>;
  using PY = PlanOf<NY, false>;
  __shared__ float2 P[G::P_FLOAT2];
  int q, p;
  plane_decode<2>(&q, &p);
  if (q >= a.nz + (a.kernel ? a.kz : 0)) return;  // workgroup-uniform, before any barrier
  const bool image = q < a.nz;
  const int plane = image ? q : q - a.nz;
  const int c = threadIdx.x / G::TX, t0 = threadIdx.x % G::TX;
  const bool lane = c < G::XC;
  const int t = lane ? t0 : 0;
  const int ncols = G::NC - p;
  const float* img = image ? a.src + (long)plane * NX * NY : nullptr;
  const bool vec = image && (reinterpret_cast<uintptr_t>(img) & 15) == 0;
  auto pre = [&](float2 z0, float2 z1, int n) {
    return p ? cmul(csub(z0, z1), a.fx.tw[n]) : cadd(z0, z1);
  };
  for (int g0 = 0; g0 < G::PAIRS; g0 += G::XC) {
    const int pr = g0 + c;
    const bool act = lane && pr < G::PAIRS;
    const int ya = 2 * pr, yb = ya + 1;
    float2* E = P + ya * G::PB;  // the pair's two rows
    // u_p[n], n < H, into the pair's rows at the exchange positions
    if (act) {
      if (vec) {
        const float4* ra = reinterpret_cast<const float4*>(img + (long)ya * NX);
        const float4* rb = reinterpret_cast<const float4*>(img + (long)yb * NX);
        for (int f = t; f < G::H / 4; f += G::TX) {
          const float4 a0 = ra[f], a1 = ra[f + G::H / 4], b0 = rb[f], b1 = rb[f + G::H / 4];
          const int n = 4 * f;
          E[XRow<G::H>::at(n)] = pre(make_float2(a0.x, b0.x), make_float2(a1.x, b1.x), n);
          E[XRow<G::H>::at(n + 1)] = pre(make_float2(a0.y, b0.y), make_float2(a1.y, b1.y), n + 1);
          E[XRow<G::H>::at(n + 2)] = pre(make_float2(a0.z, b0.z), make_float2(a1.z, b1.z), n + 2);
          E[XRow<G::H>::at(n + 3)] = pre(make_float2(a0.w, b0.w), make_float2(a1.w, b1.w), n + 3);
        }
      } else if (image) {
        const float* ra = img + (long)ya * NX;
        const float* rb = img + (long)yb * NX;
        for (int n = t; n < G::H; n += G::TX)
          E[XRow<G::H>::at(n)] = pre(make_float2(ra[n], rb[n]), make_float2(ra[n + G::H], rb[n + G::H]), n);
      } else {
        // row y of kernel plane `plane` holds kernel row (y + cy) mod NY,
        // and its position x tap (x + cx) mod NX
        int ia = ya + a.cy, ib = yb + a.cy;
        if (ia >= NY) ia -= NY;
        if (ib >= NY) ib -= NY;
        const bool oka = ia < a.ky, okb = ib < a.ky;
        const float* ka = a.kernel + ((long)plane * a.ky + (oka ? ia : 0)) * a.kx;
        const float* kb = a.kernel + ((long)plane * a.ky + (okb ? ib : 0)) * a.kx;
        const int cx = a.cx, kx = a.kx;
        auto tap = [&](int pos) {
          int x = pos + cx;
          if (x >= NX) x -= NX;
          float xa = 0.f, xb = 0.f;
          if (x < kx) {
            if (oka) xa = ka[x];
            if (okb) xb = kb[x];
          }
          return make_float2(xa, xb);
        };
        for (int n = t; n < G::H; n += G::TX) E[XRow<G::H>::at(n)] = pre(tap(n), tap(n + G::H), n);
      }
    }
    __syncthreads();
    auto addr = [&](int pos) { return act ? ya * G::PB + XRow<G::H>::at(pos) : G::P_FLOAT2 - 1; };
    float2 v[PX::MAXE()];
    sfft_for_inputs<G::H, false>(t, [&](int slot, int pos) { v[slot] = P[addr(pos)]; });
    __syncthreads();  // every input is read before the in-place exchange
    sfft_run<G::H, false, -1>(v, t, P, addr, a.hx_stw);
    sfft_for_outputs<G::H, false>(t, [&](int slot, int pos) { P[addr(pos)] = v[slot]; });
    __syncthreads();
    // local column j is bin k = 2j + p; bin nx - k is output jm
    float2 ea[G::JX], eb[G::JX];
#pragma unroll
    for (int i = 0; i < G::JX; ++i) {
      const int j = t + i * G::TX;
      if (act && j < ncols) {
        const int jm = p ? G::H - 1 - j : (j == 0 ? 0 : G::H - j);
        const float2 Z = E[XRow<G::H>::at(j)];
        const float2 Zm = E[XRow<G::H>::at(jm)];
        ea[i] = make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
        eb[i] = make_float2(0.5f * (Z.y + Zm.y), 0.5f * (Zm.x - Z.x));
      }
    }
    __syncthreads();  // the exchange is read before its rows take the bins
#pragma unroll
    for (int i = 0; i < G::JX; ++i) {
      const int j = t + i * G::TX;
      if (act && j < ncols) {
        P[ya * G::PB + j] = ea[i];
        P[yb * G::PB + j] = eb[i];
      }
    }
    __syncthreads();
  }
  // columns: transform col of part 0 packs local columns 0 and H/2 (bins 0
  // and nx/2, both real) into col 0
  float2* out = (image ? a.spec : a.kplanes) + (long)plane * NY * a.nxp;
  const long nxp = a.nxp;
  const int cc = threadIdx.x % G::YC, u0 = threadIdx.x / G::YC;
  const bool ylane = u0 < G::TY;
  const int u = ylane ? u0 : 0;
  for (int g0 = 0; g0 < G::NCT; g0 += G::YC) {
    const int col = g0 + cc;
    const bool act = ylane && col < G::NCT;
    const bool packed = act && p == 0 && col == 0;
    auto addr = [&](int pos) { return act ? pos * G::PB + col : G::P_FLOAT2 - 1; };
    float2 v[PY::MAXE()];
    sfft_for_inputs<NY, false>(u, [&](int slot, int pos) {
      float2 x = P[addr(pos)];
      if (packed) x = make_float2(x.x, P[pos * G::PB + G::H / 2].x);
      v[slot] = x;
    });
    __syncthreads();  // every input of the column is read before the exchange
    sfft_run<NY, false, -1>(v, u, P, addr, a.fy.stw);
    if (act && !packed)
      sfft_for_outputs<NY, false>(u, [&](int slot, int pos) { out[(long)pos * nxp + 2 * col + p] = v[slot]; });
    if (p == 0 && g0 == 0) {  // workgroup-uniform
      if (packed) sfft_for_outputs<NY, false>(u, [&](int slot, int pos) { P[pos * G::PB] = v[slot]; });
      __syncthreads();
      if (packed)
        sfft_for_outputs<NY, false>(u, [&](int slot, int pos) {
          const float2 F = v[slot], Fm = P[(pos == 0 ? 0 : NY - pos) * G::PB];
          out[(long)pos * nxp] = make_float2(0.5f * (F.x + Fm.x), 0.5f * (F.y - Fm.y));
          out[(long)pos * nxp + G::H] = make_float2(0.5f * (F.y + Fm.y), 0.5f * (Fm.x - F.x));
        });
    }
  }
}

template <int NX, int NY>
struct PlaneGeomI2W {
  static constexpr int want(int a, int b) { return a > b ? a : b; }
  static constexpr int least(int a, int b) { return a < b ? a : b; }
  static constexpr int GY = NY / 2;                       // column IFFT length
  static constexpr int TX = threads_per_transform<NX>();  // per row pair
  static constexpr int TY = threads_per_transform<GY>();
  static constexpr int NB = NX / 2 + 1;
  static constexpr int NBT = NB - 1;  // column transforms (0 and nx/2 packed)
  static constexpr int LPX = XRow<NX>::LP;
  static constexpr int PB = want(NB, (LPX + 1) / 2);  // two rows hold a pair's exchange
  static constexpr int SL = 2 * PB;                    // real staging pitch (floats)
  static constexpr int PAIRS = GY / 2;
  static constexpr int NT_ = (want(TX * PAIRS, TY * NBT) + 63) / 64 * 64;
  static constexpr int NT = NT_ > 1024 ? 1024 : NT_;
  static constexpr int XC = least(NT / TX, PAIRS);
  static constexpr int YC = least(NT / TY, NBT);
  static constexpr int P_FLOAT2 = GY * PB + 1;  // + dummy slot
  static constexpr long LDS_BYTES = 8L * P_FLOAT2;
  static_assert(NX % 4 == 0 && NY % 4 == 0, "wide split inverse takes nx % 4 == 0, ny % 4 == 0");
  static_assert(XC >= 1 && YC >= 1, "plane too wide for one workgroup");
  static_assert(LDS_BYTES <= 160 * 1024, "half plane does not fit in LDS");
};

template <int NX, int NY>
__global__ void __launch_bounds__((PlaneGeomI2W<NX, NY>::NT)) sk_plane_inv2w(PlaneArgs a) {
  using G = PlaneGeomI2W<NX, NY>;
  using PX = PlanOf<NX, false>;
  using PY = PlanOf<G::GY, false>;
  __shared__ float2 P[G::P_FLOAT2];
  int plane, q;
  plane_decode<2>(&plane, &q);
  if (plane >= a.nz) return;  // workgroup-uniform, before any barrier
  {
    // row k of P: X[k] + X[k + g] (part 0) or (X[k] - X[k + g]) W_ny^-k
    const float2* sp = a.spec + (long)plane * NY * a.nxp;
    const long half = (long)G::GY * a.nxp;
    constexpr int N = G::GY * G::NB, IT = (N + G::NT - 1) / G::NT, CH = 4;
    for (int j0 = 0; j0 < IT; j0 += CH) {
      float2 r0[CH], r1[CH], w[CH];
#pragma unroll
      for (int j = 0; j < CH; ++j) {
        const int i = threadIdx.x + (j0 + j) * G::NT;
        const int k = i / G::NB;
        const long o = (long)k * a.nxp + (i - k * G::NB);
        r0[j] = i < N ? sp[o] : make_float2(0.f, 0.f);
        r1[j] = i < N ? sp[o + half] : make_float2(0.f, 0.f);
        w[j] = q && i < N ? a.fy.tw[k] : make_float2(1.f, 0.f);
      }
#pragma unroll
      for (int j = 0; j < CH; ++j) {
        const int i = threadIdx.x + (j0 + j) * G::NT;
        const int k = i / G::NB;
        if (i < N)
          P[k * G::PB + (i - k * G::NB)] = q ? cmulc(csub(r0[j], r1[j]), w[j]) : cadd(r0[j], r1[j]);
      }
    }
  }
  __syncthreads();
  {
    // the columns' length-g IFFTs, in place; transform 0 packs columns 0
    // (real part) and nx/2 (imaginary part)
    const int cc = threadIdx.x % G::YC, u0 = threadIdx.x / G::YC;
    const bool ylane = u0 < G::TY;
    const int u = ylane ? u0 : 0;
    for (int g0 = 0; g0 < G::NBT; g0 += G::YC) {
      const int col = g0 + cc;
      const bool act = ylane && col < G::NBT;
      const bool packed = act && col == 0;
      auto addr = [&](int pos) { return act ? pos * G::PB + col : G::P_FLOAT2 - 1; };
      float2 v[PY::MAXE()];
      sfft_for_inputs<G::GY, false>(u, [&](int slot, int pos) {
        float2 x = P[addr(pos)];
        if (packed) {
          const float2 b = P[pos * G::PB + NX / 2];
          x = make_float2(x.x - b.y, x.y + b.x);
        }
        v[slot] = x;
      });
      __syncthreads();  // every input of the column is read before the exchange
      sfft_run<G::GY, false, +1>(v, u, P, addr, a.hy_stw);
      if (packed) {
        sfft_for_outputs<G::GY, false>(u, [&](int slot, int pos) {
          P[pos * G::PB] = make_float2(v[slot].x, 0.f);
          P[pos * G::PB + NX / 2] = make_float2(v[slot].y, 0.f);
        });
      } else {
        sfft_for_outputs<G::GY, false>(u, [&](int slot, int pos) { P[addr(pos)] = v[slot]; });
      }
    }
  }
  __syncthreads();
  // x-C2R of the part's row pairs, exchanging in the pair's own rows; the
  // real rows go to the staging pitch inside them
  float* st = reinterpret_cast<float*>(P);
  const int c = threadIdx.x / G::TX, t0 = threadIdx.x % G::TX;
  const bool lane = c < G::XC;
  const int t = lane ? t0 : 0;
  for (int g0 = 0; g0 < G::PAIRS; g0 += G::XC) {
    const int pr = g0 + c;
    const bool act = lane && pr < G::PAIRS;
    const int ya = 2 * pr, yb = ya + 1;
    float2 v[PX::MAXE()];
    sfft_for_inputs<NX, false>(t, [&](int slot, int pos) {
      const bool mirror = pos >= G::NB;
      const int kb = mirror ? NX - pos : pos;
      float2 A = act ? P[ya * G::PB + kb] : make_float2(0.f, 0.f);
      float2 B = act ? P[yb * G::PB + kb] : make_float2(0.f, 0.f);
      if (kb == 0 || 2 * kb == NX) {
        A.y = 0.f;
        B.y = 0.f;
      }
      if (mirror) {
        A.y = -A.y;
        B.y = -B.y;
      }
      v[slot] = make_float2(A.x - B.y, A.y + B.x);
    });
    __syncthreads();  // every input is read before the in-place exchange
    auto addr = [&](int pos) { return act ? ya * G::PB + XRow<NX>::at(pos) : G::P_FLOAT2 - 1; };
    sfft_run<NX, false, +1>(v, t, P, addr, a.fx.stw);
    if (act)
      sfft_for_outputs<NX, false>(t, [&](int slot, int pos) {
        st[ya * G::SL + pos] = v[slot].x;
        st[yb * G::SL + pos] = v[slot].y;
      });
    __syncthreads();
  }
  // local row m is image row 2m + q
  float* g = a.dst + (long)plane * NX * NY + (long)q * NX;
  if ((reinterpret_cast<uintptr_t>(g) & 15) == 0) {
    constexpr int Q = NX / 4, N4 = Q * G::GY, IT = (N4 + G::NT - 1) / G::NT;
#pragma unroll
    for (int j = 0; j < IT; ++j) {
      const int i = threadIdx.x + j * G::NT;
      if (i < N4) {
        const int m = i / Q, x4 = i - m * Q;
        *reinterpret_cast<float4*>(g + (long)2 * m * NX + 4 * x4) =
            *reinterpret_cast<const float4*>(st + m * G::SL + 4 * x4);
      }
    }
  } else {
    for (int i = threadIdx.x; i < NX * G::GY; i += G::NT) {
      const int m = i / NX, x = i - m * NX;
      g[(long)2 * m * NX + x] = st[m * G::SL + x];
    }
  }
}

template <int NX, int NY>
void plane_launch_fwd(const PlaneArgs& a, hipStream_t st) {
  using G = PlaneGeom<NX, NY>;
  const int nblk = a.nz + (a.kernel ? a.kz : 0);
  hipLaunchKernelGGL((sk_plane_fwd<NX, NY>), dim3((unsigned)nblk), dim3(G::NT), 0, st, a);
}

template <int NX, int NY>
void plane_launch_inv(const PlaneArgs& a, hipStream_t st) {
  using G = PlaneGeom<NX, NY>;
  hipLaunchKernelGGL((sk_plane_inv<NX, NY>), dim3((unsigned)a.nz), dim3(G::NT), 0, st, a);
}

// the split kernels need compile-time plans of the half lengths; the
// forward split measured neutral or slower for planes below 10240 points
// (nx x ny x nz, GPU time per convolution: 64 x 64 x 64 15.19 -> 15.12 us,
// 96 x 96 x 64 22.59 -> 23.44 us, 64 x 128 x 64 19.41 -> 19.24 us) and
// faster from there (128 x 80 x 64: 23.27 -> 22.26 us, 128 x 128 x 64:
// 28.58 -> 27.10 us, 256 x 64 x 32: 26.05 -> 24.44 us, 64 x 256 x 32:
// 23.95 -> 22.79 us)
template <int NX, int NY>
constexpr bool plane_fwd2_ok() {
  return NX * NY >= 10240 && static_plan(NX / 2).n > 0;
}
template <int NX, int NY>
constexpr bool plane_inv2_ok() {
  return NY % 4 == 0 && static_plan(NY / 2).n > 0;
}

template <int NX, int NY>
void plane_launch_fwd2(const PlaneArgs& a, hipStream_t st) {
  using G = PlaneGeomF2<NX, NY>;
  const int nblk = plane_blocks<2>(a.nz + (a.kernel ? a.kz : 0));
  hipLaunchKernelGGL((sk_plane_fwd2<NX, NY>), dim3((unsigned)nblk), dim3(G::NT), 0, st, a);
}

template <int NX, int NY>
void plane_launch_inv2(const PlaneArgs& a, hipStream_t st) {
  using G = PlaneGeomI2<NX, NY>;
  hipLaunchKernelGGL((sk_plane_inv2<NX, NY>), dim3((unsigned)plane_blocks<2>(a.nz)), dim3(G::NT),
                     0, st, a);
}

// workgroups of a wide split kernel one CU holds at once (its LDS, threads
// and registers), queried once
template <int NX, int NY>
int plane_occupancy_fwd2w() {
  static const int n = [] {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, sk_plane_fwd2w<NX, NY>,
                                                     PlaneGeomF2W<NX, NY>::NT, 0) != hipSuccess)
      b = 1;
    return b > 0 ? b : 1;
  }();
  return n;
}
template <int NX, int NY>
int plane_occupancy_inv2w() {
  static const int n = [] {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, sk_plane_inv2w<NX, NY>,
                                                     PlaneGeomI2W<NX, NY>::NT, 0) != hipSuccess)
      b = 1;
    return b > 0 ? b : 1;
  }();
  return n;
}

template <int NX, int NY>
void plane_launch_fwd2w(const PlaneArgs& a, hipStream_t st) {
  using G = PlaneGeomF2W<NX, NY>;
  const int nblk = plane_blocks<2>(a.nz + (a.kernel ? a.kz : 0));
  hipLaunchKernelGGL((sk_plane_fwd2w<NX, NY>), dim3((unsigned)nblk), dim3(G::NT), 0, st, a);
}

template <int NX, int NY>
void plane_launch_inv2w(const PlaneArgs& a, hipStream_t st) {
  using G = PlaneGeomI2W<NX, NY>;
  hipLaunchKernelGGL((sk_plane_inv2w<NX, NY>), dim3((unsigned)plane_blocks<2>(a.nz)), dim3(G::NT),
                     0, st, a);
}

}  // namespace fcamd
