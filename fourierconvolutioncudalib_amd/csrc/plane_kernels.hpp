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

}  // namespace fcamd
