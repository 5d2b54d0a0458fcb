// fft_radix.hpp -- complex helpers and in-register DFT butterflies.
//
// DIR = -1 is the forward transform X[k] = sum_n x[n] exp(-2*pi*i*n*k/R)
// (the sign cuFFT's R2C uses in src/convolution3Dfft.cu:523-525), DIR = +1
// the unnormalised inverse (C2R at :547).  Constants are exact values of
// cos/sin at rational multiples of pi, written out to double precision and
// rounded once to float.
#pragma once

#include <hip/hip_runtime.h>

namespace fcamd {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// a * (S * i), S = +-1
template <int S>
__device__ __forceinline__ float2 rot90(float2 a) {
  return S > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}
// twiddle exp(DIR*2*pi*i*e/L) from the forward table
template <int DIR>
__device__ __forceinline__ float2 twiddle(const float2* tw, int e) {
  float2 w = tw[e];
  return DIR < 0 ? w : cconj(w);
}

namespace k {
constexpr float kS3 = 0.86602540378443864676f;   // sin(2pi/3)
constexpr float kC51 = 0.30901699437494742410f;  // cos(2pi/5)
constexpr float kC52 = -0.80901699437494742410f; // cos(4pi/5)
constexpr float kS51 = 0.95105651629515357212f;  // sin(2pi/5)
constexpr float kS52 = 0.58778525229247312917f;  // sin(4pi/5)
constexpr float kC71 = 0.62348980185873353053f;  // cos(2pi/7)
constexpr float kC72 = -0.22252093395631440429f; // cos(4pi/7)
constexpr float kC73 = -0.90096886790241912624f; // cos(6pi/7)
constexpr float kS71 = 0.78183148246802980871f;  // sin(2pi/7)
constexpr float kS72 = 0.97492791218182360702f;  // sin(4pi/7)
constexpr float kS73 = 0.43388373911755812048f;  // sin(6pi/7)
constexpr float kR2 = 0.70710678118654752440f;   // sqrt(1/2)
}  // namespace k

// ---------------------------------------------------------------------------
// compile-time sin/cos of 2*pi*m/n in double precision (range reduction by
// symmetry to [0, pi/4], then Taylor series), rounded once to float
// ---------------------------------------------------------------------------
namespace cx {
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double sin_series(double x) {
  double term = x, sum = x;
  for (int k = 1; k < 14; ++k) {
    term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double cos_series(double x) {
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 14; ++k) {
    term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}
// cos and sin of 2*pi*m/n for integer m, n > 0, exact symmetry handling:
// the angle is reduced in units of n/8 so the series only sees [0, pi/4].
struct CS {
  double c, s;
};
constexpr CS cossin_2pi(long m, long n) {
  m %= n;
  if (m < 0) m += n;
  // octant o = floor(8 m / n), remainder r in [0, n/8) as a fraction
  const long m8 = 8 * m;
  const long o = m8 / n;
  const double frac = (double)(m8 - o * n) / (double)n;  // in [0, 1)
  const double x = frac * (kPi / 4.0);                   // [0, pi/4)
  const double y = (kPi / 4.0) - x;                      // (0, pi/4]
  double c = 0.0, s = 0.0;
  switch (o) {
    case 0: c = cos_series(x); s = sin_series(x); break;
    case 1: c = sin_series(y); s = cos_series(y); break;
    case 2: c = -sin_series(x); s = cos_series(x); break;
    case 3: c = -cos_series(y); s = sin_series(y); break;
    case 4: c = -cos_series(x); s = -sin_series(x); break;
    case 5: c = -sin_series(y); s = -cos_series(y); break;
    case 6: c = sin_series(x); s = -cos_series(x); break;
    default: c = cos_series(y); s = -sin_series(y); break;
  }
  return CS{c, s};
}
template <int R>
struct Roots {
  float c[R];
  float s[R];
  constexpr Roots() : c(), s() {
    for (int m = 0; m < R; ++m) {
      const CS v = cossin_2pi(m, R);
      c[m] = (float)v.c;
      s[m] = (float)v.s;
    }
  }
};
constexpr int smallest_factor(int r) {
  for (int f : {4, 2, 3, 5, 7})
    if (r % f == 0 && r != f) return f;
  return r;
}
}  // namespace cx

// Composite radix R = R1 * R2 by an in-register Cooley-Tukey split:
// y[k1 + R1 k2] = sum_n2 W_R^(n2 k1) W_R2^(n2 k2) sum_n1 x[n1 R2 + n2] W_R1^(n1 k1).
// Radices with a dedicated butterfly below specialise this template.
template <int R, int DIR>
struct Dft {
  static constexpr int R1 = cx::smallest_factor(R);
  static constexpr int R2 = R / R1;
  static_assert(R1 > 1 && R2 > 1, "composite radix expected");
  __device__ __forceinline__ static void run(float2* v) {
    constexpr cx::Roots<R> W{};
    float2 t[R];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      float2 a[R1];
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) a[n1] = v[n1 * R2 + n2];
      Dft<R1, DIR>::run(a);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) {
        const int m = (n2 * k1) % R;
        if (m == 0) {
          t[k1 * R2 + n2] = a[k1];
        } else {
          const float2 w = make_float2(W.c[m], DIR * W.s[m]);
          t[k1 * R2 + n2] = cmul(a[k1], w);
        }
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      float2 b[R2];
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) b[n2] = t[k1 * R2 + n2];
      Dft<R2, DIR>::run(b);
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = b[k2];
    }
  }
};

template <int DIR>
struct Dft<1, DIR> {
  __device__ __forceinline__ static void run(float2*) {}
};

template <int DIR>
struct Dft<2, DIR> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <int DIR>
struct Dft<3, DIR> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 s = cadd(v[1], v[2]);
    float2 d = csub(v[1], v[2]);
    float2 t = make_float2(v[0].x - 0.5f * s.x, v[0].y - 0.5f * s.y);
    // i * DIR * sin(2pi/3) * d
    float2 u = rot90<DIR>(cscale(d, k::kS3));
    v[0] = cadd(v[0], s);
    v[1] = cadd(t, u);
    v[2] = csub(t, u);
  }
};

template <int DIR>
struct Dft<4, DIR> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    float2 c = cadd(v[1], v[3]), d = rot90<DIR>(csub(v[1], v[3]));
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, d);
    v[3] = csub(b, d);
  }
};

template <int DIR>
struct Dft<5, DIR> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    float2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    float2 x0 = v[0];
    float2 r1 = make_float2(x0.x + k::kC51 * a1.x + k::kC52 * a2.x, x0.y + k::kC51 * a1.y + k::kC52 * a2.y);
    float2 r2 = make_float2(x0.x + k::kC52 * a1.x + k::kC51 * a2.x, x0.y + k::kC52 * a1.y + k::kC51 * a2.y);
    float2 i1 = rot90<DIR>(make_float2(k::kS51 * b1.x + k::kS52 * b2.x, k::kS51 * b1.y + k::kS52 * b2.y));
    float2 i2 = rot90<DIR>(make_float2(k::kS52 * b1.x - k::kS51 * b2.x, k::kS52 * b1.y - k::kS51 * b2.y));
    v[0] = make_float2(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
    v[1] = cadd(r1, i1);
    v[4] = csub(r1, i1);
    v[2] = cadd(r2, i2);
    v[3] = csub(r2, i2);
  }
};

template <int DIR>
struct Dft<7, DIR> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a1 = cadd(v[1], v[6]), b1 = csub(v[1], v[6]);
    float2 a2 = cadd(v[2], v[5]), b2 = csub(v[2], v[5]);
    float2 a3 = cadd(v[3], v[4]), b3 = csub(v[3], v[4]);
    float2 x0 = v[0];
    // y_k = x0 + sum_n a_n cos(2pi nk/7) + DIR*i sum_n b_n sin(2pi nk/7)
    float2 r1 = make_float2(x0.x + k::kC71 * a1.x + k::kC72 * a2.x + k::kC73 * a3.x,
                            x0.y + k::kC71 * a1.y + k::kC72 * a2.y + k::kC73 * a3.y);
    float2 r2 = make_float2(x0.x + k::kC72 * a1.x + k::kC73 * a2.x + k::kC71 * a3.x,
                            x0.y + k::kC72 * a1.y + k::kC73 * a2.y + k::kC71 * a3.y);
    float2 r3 = make_float2(x0.x + k::kC73 * a1.x + k::kC71 * a2.x + k::kC72 * a3.x,
                            x0.y + k::kC73 * a1.y + k::kC71 * a2.y + k::kC72 * a3.y);
    // sin(2pi*n*k/7) for k=1: s1 s2 s3; k=2: s2 -s3 -s1; k=3: s3 -s1 s2
    float2 i1 = rot90<DIR>(make_float2(k::kS71 * b1.x + k::kS72 * b2.x + k::kS73 * b3.x,
                                       k::kS71 * b1.y + k::kS72 * b2.y + k::kS73 * b3.y));
    float2 i2 = rot90<DIR>(make_float2(k::kS72 * b1.x - k::kS73 * b2.x - k::kS71 * b3.x,
                                       k::kS72 * b1.y - k::kS73 * b2.y - k::kS71 * b3.y));
    float2 i3 = rot90<DIR>(make_float2(k::kS73 * b1.x - k::kS71 * b2.x + k::kS72 * b3.x,
                                       k::kS73 * b1.y - k::kS71 * b2.y + k::kS72 * b3.y));
    v[0] = make_float2(x0.x + a1.x + a2.x + a3.x, x0.y + a1.y + a2.y + a3.y);
    v[1] = cadd(r1, i1);
    v[6] = csub(r1, i1);
    v[2] = cadd(r2, i2);
    v[5] = csub(r2, i2);
    v[3] = cadd(r3, i3);
    v[4] = csub(r3, i3);
  }
};

template <int DIR>
struct Dft<8, DIR> {
  __device__ __forceinline__ static void run(float2* v) {
    // radix-2 split into two 4-point DFTs of the even and odd samples
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, DIR>::run(e);
    Dft<4, DIR>::run(o);
    // W8^1 = sqrt(1/2) * (1 + DIR*i), W8^2 = DIR*i, W8^3 = sqrt(1/2)*(-1 + DIR*i)
    float2 o1 = make_float2(k::kR2 * (o[1].x - DIR * o[1].y), k::kR2 * (o[1].y + DIR * o[1].x));
    float2 o2 = rot90<DIR>(o[2]);
    float2 o3 = make_float2(k::kR2 * (-o[3].x - DIR * o[3].y), k::kR2 * (-o[3].y + DIR * o[3].x));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
};

template <int DIR>
struct Dft<16, DIR> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 e[8], o[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      e[i] = v[2 * i];
      o[i] = v[2 * i + 1];
    }
    Dft<8, DIR>::run(e);
    Dft<8, DIR>::run(o);
    // W16^j = cos(pi j/8) + DIR*i*sin(pi j/8)
    constexpr float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
    const float cs[8] = {1.f, c1, k::kR2, s1, 0.f, -s1, -k::kR2, -c1};
    const float sn[8] = {0.f, s1, k::kR2, c1, 1.f, c1, k::kR2, s1};
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      float2 w = make_float2(cs[j], DIR * sn[j]);
      float2 t = cmul(o[j], w);
      v[j] = cadd(e[j], t);
      v[j + 8] = csub(e[j], t);
    }
  }
};

// Odd prime radix R as a direct DFT over the pair sums and differences
// a_n = x_n + x_(R-n), b_n = x_n - x_(R-n), n = 1..(R-1)/2:
//   y_k = x_0 + sum_n a_n cos(2 pi n k / R) + DIR i sum_n b_n sin(2 pi n k / R)
// and y_(R-k) the same with the sine sum subtracted (the radix-7 form above
// for any odd prime).
template <int R, int DIR>
struct PrimeDft {
  __device__ __forceinline__ static void run(float2* v) {
    constexpr int H = (R - 1) / 2;
    constexpr cx::Roots<R> W{};
    float2 a[H + 1], b[H + 1];
    const float2 x0 = v[0];
    float2 sum = x0;
#pragma unroll
    for (int n = 1; n <= H; ++n) {
      a[n] = cadd(v[n], v[R - n]);
      b[n] = csub(v[n], v[R - n]);
      sum = cadd(sum, a[n]);
    }
#pragma unroll
    for (int k = 1; k <= H; ++k) {
      float2 re = x0, im = make_float2(0.f, 0.f);
#pragma unroll
      for (int n = 1; n <= H; ++n) {
        const int m = (n * k) % R;
        re.x += W.c[m] * a[n].x;
        re.y += W.c[m] * a[n].y;
        im.x += W.s[m] * b[n].x;
        im.y += W.s[m] * b[n].y;
      }
      const float2 i = rot90<DIR>(im);
      v[k] = cadd(re, i);
      v[R - k] = csub(re, i);
    }
    v[0] = sum;
  }
};

template <int DIR>
struct Dft<11, DIR> : PrimeDft<11, DIR> {};
template <int DIR>
struct Dft<13, DIR> : PrimeDft<13, DIR> {};

}  // namespace fcamd
