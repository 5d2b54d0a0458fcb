/* fc_oracle.c -- CPU oracle for the 3D FFT convolution hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see fc_oracle.h).  Plain C99, FP64 except where a
 * function restates float arithmetic of the reference bit for bit.
 * Citations are relative to the reference repository root.
 */
#include "fc_oracle.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#include <math.h>
#include <stdlib.h>
#include <string.h>

static long vol(const int* s) { return (long)s[0] * s[1] * s[2]; }

/* ------------------------------------------------------------------------ */
/* tests/test_algorithms.hpp:9-58                                            */
/* ------------------------------------------------------------------------ */
void fco_convolve_host(const float* image, const int* ishape,
                       const float* kernel, const int* kshape, float* result,
                       const int* off) {
  if (vol(ishape) == 0) return;                       /* :13 */
  const int hk0 = kshape[0] / 2, hk1 = kshape[1] / 2, hk2 = kshape[2] / 2;
  const long sy = ishape[2], sz = (long)ishape[1] * ishape[2];
  const long ky_s = kshape[2], kz_s = (long)kshape[1] * kshape[2];
  for (int iz = off[0]; iz < ishape[0] - off[0]; ++iz) {       /* :22 */
    for (int iy = off[1]; iy < ishape[1] - off[1]; ++iy) {     /* :24 */
      for (int ix = off[2]; ix < ishape[2] - off[2]; ++ix) {   /* :26 */
        float value = 0.f;                                     /* :33 */
        for (int kz = 0; kz < kshape[0]; ++kz) {               /* :37 */
          for (int ky = 0; ky < kshape[1]; ++ky) {             /* :38 */
            for (int kx = 0; kx < kshape[2]; ++kx) {           /* :40 */
              /* :43-45 the kernel is read flipped */
              float kernel_value =
                  kernel[(kshape[0] - 1 - kz) * kz_s +
                         (kshape[1] - 1 - ky) * ky_s + (kshape[2] - 1 - kx)];
              /* :46-48 */
              float image_value =
                  image[(iz - hk0 + kz) * sz + (iy - hk1 + ky) * sy +
                        (ix - hk2 + kx)];
              value += kernel_value * image_value;             /* :50 */
            }
          }
        }
        result[iz * sz + iy * sy + ix] = value;                /* :54 */
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* src/convolution3Dfft.cu:128-166 (placement of the kernel taps)            */
/* ------------------------------------------------------------------------ */
int fco_fftshift_pad(const float* kernel, const int* kshape, double* padded,
                     const int* ishape) {
  for (int a = 0; a < 3; ++a)
    if (kshape[a] > ishape[a] || kshape[a] < 1) return -1;
  memset(padded, 0, sizeof(double) * (size_t)vol(ishape));
  for (int z = 0; z < kshape[0]; ++z)
    for (int y = 0; y < kshape[1]; ++y)
      for (int x = 0; x < kshape[2]; ++x) {
        /* :150-158 centre at k/2 then wrap negative offsets */
        int pz = z - kshape[0] / 2, py = y - kshape[1] / 2,
            px = x - kshape[2] / 2;
        if (pz < 0) pz += ishape[0];
        if (py < 0) py += ishape[1];
        if (px < 0) px += ishape[2];
        padded[((long)pz * ishape[1] + py) * ishape[2] + px] =
            kernel[((long)z * kshape[1] + y) * kshape[2] + x];
      }
  return 0;
}

int fco_circular_conv_direct(const float* image, const int* s,
                             const float* kernel, const int* k, double* out) {
  for (int a = 0; a < 3; ++a)
    if (k[a] > s[a] || k[a] < 1) return -1;
  const int c0 = k[0] / 2, c1 = k[1] / 2, c2 = k[2] / 2;
  for (int iz = 0; iz < s[0]; ++iz)
    for (int iy = 0; iy < s[1]; ++iy)
      for (int ix = 0; ix < s[2]; ++ix) {
        double acc = 0.0;
        for (int jz = 0; jz < k[0]; ++jz) {
          int z = ((iz - (jz - c0)) % s[0] + s[0]) % s[0];
          for (int jy = 0; jy < k[1]; ++jy) {
            int y = ((iy - (jy - c1)) % s[1] + s[1]) % s[1];
            for (int jx = 0; jx < k[2]; ++jx) {
              int x = ((ix - (jx - c2)) % s[2] + s[2]) % s[2];
              acc += (double)kernel[((long)jz * k[1] + jy) * k[2] + jx] *
                     (double)image[((long)z * s[1] + y) * s[2] + x];
            }
          }
        }
        out[((long)iz * s[1] + iy) * s[2] + ix] = acc;
      }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Boundary-mode convolution, direct FP64 sum.  Same kernel orientation and   */
/* centring as fco_circular_conv_direct (taps at offset j - k/2, flipped as  */
/* in tests/test_algorithms.hpp:40-47); image positions outside [0, n) are   */
/* mapped by the boundary rule instead of wrapping.  The rules are those of  */
/* the header's FCAMD_BOUNDARY_* (scipy.ndimage 'wrap', 'constant' with 0,   */
/* 'nearest', 'reflect', 'mirror').  No counterpart in the reference, whose  */
/* header leaves boundary conditions as a TODO (src/convolution3Dfft.h:39).  */
/* ------------------------------------------------------------------------ */
static int bmap(int i, int n, int mode, int* ok) {
  *ok = 1;
  switch (mode) {
    case 1: /* zero */
      *ok = i >= 0 && i < n;
      return *ok ? i : 0;
    case 2: /* nearest */
      return i < 0 ? 0 : (i >= n ? n - 1 : i);
    case 3: { /* reflect, period 2n */
      int r = i % (2 * n);
      if (r < 0) r += 2 * n;
      return r < n ? r : 2 * n - 1 - r;
    }
    case 4: { /* mirror, period 2n - 2 */
      int per, r;
      if (n == 1) return 0;
      per = 2 * n - 2;
      r = i % per;
      if (r < 0) r += per;
      return r < n ? r : per - r;
    }
    default: { /* periodic */
      int r = i % n;
      if (r < 0) r += n;
      return r;
    }
  }
}

int fco_boundary_conv_direct(const float* image, const int* s, const float* kernel,
                             const int* k, int mode, int correlate, double* out) {
  int a;
  const int c0 = k[0] / 2, c1 = k[1] / 2, c2 = k[2] / 2;
  if (mode < 0 || mode > 4) return -1;
  for (a = 0; a < 3; ++a)
    if (k[a] < 1 || s[a] < 1) return -1;
  for (int iz = 0; iz < s[0]; ++iz)
    for (int iy = 0; iy < s[1]; ++iy)
      for (int ix = 0; ix < s[2]; ++ix) {
        double acc = 0.0;
        for (int jz = 0; jz < k[0]; ++jz) {
          int okz, oky, okx;
          /* convolution reads i - (j - c), correlation i + (j - c) */
          const int dz = correlate ? jz - c0 : c0 - jz;
          const int z = bmap(iz + dz, s[0], mode, &okz);
          if (!okz) continue;
          for (int jy = 0; jy < k[1]; ++jy) {
            const int dy = correlate ? jy - c1 : c1 - jy;
            const int y = bmap(iy + dy, s[1], mode, &oky);
            if (!oky) continue;
            for (int jx = 0; jx < k[2]; ++jx) {
              const int dx = correlate ? jx - c2 : c2 - jx;
              const int x = bmap(ix + dx, s[2], mode, &okx);
              if (!okx) continue;
              acc += (double)kernel[((long)jz * k[1] + jy) * k[2] + jx] *
                     (double)image[((long)z * s[1] + y) * s[2] + x];
            }
          }
        }
        out[((long)iz * s[1] + iy) * s[2] + ix] = acc;
      }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* FP64 DFT of any length: recursive decimation in time over prime factors.  */
/* Stands in for cuFFT (closed library; see src/convolution3Dfft.cu:519-547) */
/* by the textbook definition X[k] = sum_j x[j] exp(sign*2*pi*i*j*k/n).      */
/* ------------------------------------------------------------------------ */
typedef struct {
  int n;
  int nf;
  int f[40];
  double* tw; /* exp(-2*pi*i*j/n), j < n, interleaved */
  double* tmp;
} fplan;

static void fplan_init(fplan* p, int n) {
  p->n = n;
  p->nf = 0;
  int m = n;
  for (int d = 2; (long)d * d <= m; ++d)
    while (m % d == 0) { p->f[p->nf++] = d; m /= d; }
  if (m > 1) p->f[p->nf++] = m;
  p->tw = (double*)malloc(sizeof(double) * 2 * (size_t)n);
  for (int j = 0; j < n; ++j) {
    /* angle reduced to [0, 2pi) by the exact integer j < n */
    double a = -2.0 * M_PI * (double)j / (double)n;
    p->tw[2 * j] = cos(a);
    p->tw[2 * j + 1] = sin(a);
  }
  p->tmp = (double*)malloc(sizeof(double) * 2 * (size_t)(n > 1 ? n : 1) * 2);
}

static void fplan_free(fplan* p) {
  free(p->tw);
  free(p->tmp);
}

/* out[0..n) <- DFT_n of in[0], in[is], ..., in[(n-1)*is] (complex) */
static void fft_rec(const fplan* P, const double* in, long is, double* out,
                    int n, int fi, int sign, double* scratch) {
  if (n == 1) {
    out[0] = in[0];
    out[1] = in[1];
    return;
  }
  const int p = P->f[fi];
  const int m = n / p;
  for (int r = 0; r < p; ++r)
    fft_rec(P, in + 2 * r * is, is * p, out + 2 * (long)r * m, m, fi + 1, sign,
            scratch);
  const long step = P->n / n; /* W_n^e = W_N^(e*step) */
  for (int k = 0; k < m; ++k) {
    for (int s = 0; s < p; ++s) {
      double ar = 0.0, ai = 0.0;
      const long kk = k + (long)s * m;
      for (int r = 0; r < p; ++r) {
        long e = ((long)r * kk) % n;
        double wr = P->tw[2 * e * step], wi = P->tw[2 * e * step + 1];
        if (sign > 0) wi = -wi;
        double yr = out[2 * ((long)r * m + k)], yi = out[2 * ((long)r * m + k) + 1];
        ar += wr * yr - wi * yi;
        ai += wr * yi + wi * yr;
      }
      scratch[2 * s] = ar;
      scratch[2 * s + 1] = ai;
    }
    for (int s = 0; s < p; ++s) {
      out[2 * (k + (long)s * m)] = scratch[2 * s];
      out[2 * (k + (long)s * m) + 1] = scratch[2 * s + 1];
    }
  }
}

static void fft_line(const fplan* P, double* data, long stride, int sign,
                     double* line, double* res) {
  const int n = P->n;
  for (int i = 0; i < n; ++i) {
    line[2 * i] = data[2 * i * stride];
    line[2 * i + 1] = data[2 * i * stride + 1];
  }
  fft_rec(P, line, 1, res, n, 0, sign, P->tmp);
  for (int i = 0; i < n; ++i) {
    data[2 * i * stride] = res[2 * i];
    data[2 * i * stride + 1] = res[2 * i + 1];
  }
}

void fco_fft1d(double* data, int n, long stride, int sign) {
  if (n <= 1) return;
  fplan P;
  fplan_init(&P, n);
  double* line = (double*)malloc(sizeof(double) * 4 * (size_t)n);
  fft_line(&P, data, stride, sign, line, line + 2 * n);
  free(line);
  fplan_free(&P);
}

void fco_fft3d(double* data, const int* s, int sign) {
  const long nx = s[2], ny = s[1], nz = s[0];
  int maxn = s[0];
  if (s[1] > maxn) maxn = s[1];
  if (s[2] > maxn) maxn = s[2];
  double* line = (double*)malloc(sizeof(double) * 4 * (size_t)maxn);
  double* res = line + 2 * maxn;
  if (nx > 1) {
    fplan P;
    fplan_init(&P, (int)nx);
    for (long r = 0; r < nz * ny; ++r) fft_line(&P, data + 2 * r * nx, 1, sign, line, res);
    fplan_free(&P);
  }
  if (ny > 1) {
    fplan P;
    fplan_init(&P, (int)ny);
    for (long z = 0; z < nz; ++z)
      for (long x = 0; x < nx; ++x)
        fft_line(&P, data + 2 * (z * ny * nx + x), nx, sign, line, res);
    fplan_free(&P);
  }
  if (nz > 1) {
    fplan P;
    fplan_init(&P, (int)nz);
    for (long y = 0; y < ny; ++y)
      for (long x = 0; x < nx; ++x)
        fft_line(&P, data + 2 * (y * nx + x), nx * ny, sign, line, res);
    fplan_free(&P);
  }
  free(line);
}

/* FFT circular convolution of a double image with the shifted kernel. */
static int conv_fft_d(const double* image, const int* s, const float* kernel,
                      const int* k, double* out, int conj_kernel) {
  const long N = vol(s);
  double* pk = (double*)malloc(sizeof(double) * (size_t)N);
  if (fco_fftshift_pad(kernel, k, pk, s) != 0) {
    free(pk);
    return -1;
  }
  double* A = (double*)malloc(sizeof(double) * 2 * (size_t)N);
  double* B = (double*)malloc(sizeof(double) * 2 * (size_t)N);
  for (long i = 0; i < N; ++i) {
    A[2 * i] = image[i];
    A[2 * i + 1] = 0.0;
    B[2 * i] = pk[i];
    B[2 * i + 1] = 0.0;
  }
  free(pk);
  fco_fft3d(A, s, -1);
  fco_fft3d(B, s, -1);
  const double inv = 1.0 / (double)N; /* src/convolution3Dfft.cu:531 */
  for (long i = 0; i < N; ++i) {
    /* src/convolution3Dfft.cu:41-45: complex product, no conjugate */
    double ar = A[2 * i], ai = A[2 * i + 1];
    double br = B[2 * i], bi = conj_kernel ? -B[2 * i + 1] : B[2 * i + 1];
    A[2 * i] = (ar * br - ai * bi) * inv;
    A[2 * i + 1] = (ai * br + ar * bi) * inv;
  }
  fco_fft3d(A, s, +1);
  for (long i = 0; i < N; ++i) out[i] = A[2 * i];
  free(A);
  free(B);
  return 0;
}

int fco_fft_conv(const float* image, const int* s, const float* kernel,
                 const int* k, double* out, int conj_kernel) {
  const long N = vol(s);
  double* im = (double*)malloc(sizeof(double) * (size_t)N);
  for (long i = 0; i < N; ++i) im[i] = image[i];
  int rc = conv_fft_d(im, s, kernel, k, out, conj_kernel);
  free(im);
  return rc;
}

int fco_conv_grad_kernel(const double* g, const float* image, const int* s,
                         const int* k, double* dk) {
  for (int a = 0; a < 3; ++a)
    if (k[a] > s[a] || k[a] < 1) return -1;
  const int c0 = k[0] / 2, c1 = k[1] / 2, c2 = k[2] / 2;
  for (int jz = 0; jz < k[0]; ++jz)
    for (int jy = 0; jy < k[1]; ++jy)
      for (int jx = 0; jx < k[2]; ++jx) {
        const int dz = jz - c0, dy = jy - c1, dx = jx - c2;
        double acc = 0.0;
        for (int iz = 0; iz < s[0]; ++iz) {
          int z = ((iz - dz) % s[0] + s[0]) % s[0];
          for (int iy = 0; iy < s[1]; ++iy) {
            int y = ((iy - dy) % s[1] + s[1]) % s[1];
            const double* grow = g + ((long)iz * s[1] + iy) * s[2];
            const float* irow = image + ((long)z * s[1] + y) * s[2];
            for (int ix = 0; ix < s[2]; ++ix) {
              int x = ((ix - dx) % s[2] + s[2]) % s[2];
              acc += grow[ix] * (double)irow[x];
            }
          }
        }
        dk[((long)jz * k[1] + jy) * k[2] + jx] = acc;
      }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* tests/padd_utils.h                                                        */
/* ------------------------------------------------------------------------ */
void fco_zero_padd_extents(const int* s, const int* k, int factor, int* ext,
                           int* off) {
  for (int a = 0; a < 3; ++a) {
    ext[a] = s[a] + 2 * factor * (k[a] / 2); /* :13-23 add_minus_1 */
    off[a] = (k[a] / 2) * factor;            /* :25-38 minus_1_div_2 */
  }
}

void fco_insert_at_offsets(const float* image, const int* s, float* padded,
                           const int* ext, const int* off) {
  memset(padded, 0, sizeof(float) * (size_t)vol(ext));
  for (int z = 0; z < s[0]; ++z)
    for (int y = 0; y < s[1]; ++y)
      memcpy(padded + ((long)(z + off[0]) * ext[1] + (y + off[1])) * ext[2] + off[2],
             image + ((long)z * s[1] + y) * s[2], sizeof(float) * (size_t)s[2]);
}

/* tests/test_utils.hpp:76-88 */
double fco_l2norm_ref(const float* a, const float* b, long n) {
  double acc = 0.0;
  for (long i = 0; i < n; ++i) {
    double d = (double)(float)(b[i] - a[i]); /* diff_squared<float,double> */
    acc += d * d;
  }
  return sqrt(acc) / (double)n;
}

/* ------------------------------------------------------------------------ */
/* Multiview Richardson-Lucy (no counterpart in the reference; FP64).        */
/* ------------------------------------------------------------------------ */
int fco_rl_multiview(const float* views, int V, const int* s,
                     const float* kernels, const int* kshapes, int iters,
                     double eps, double* psi) {
  const long N = vol(s);
  double* blur = (double*)malloc(sizeof(double) * (size_t)N);
  double* ratio = (double*)malloc(sizeof(double) * (size_t)N);
  double* back = (double*)malloc(sizeof(double) * (size_t)N);
  double* acc = (double*)malloc(sizeof(double) * (size_t)N);
  for (long i = 0; i < N; ++i) {
    double m = 0.0;
    for (int v = 0; v < V; ++v) m += views[(long)v * N + i];
    psi[i] = m / V;
  }
  int rc = 0;
  for (int it = 0; it < iters && rc == 0; ++it) {
    memset(acc, 0, sizeof(double) * (size_t)N);
    const float* kp = kernels;
    for (int v = 0; v < V && rc == 0; ++v) {
      const int* k = kshapes + 3 * v;
      rc |= conv_fft_d(psi, s, kp, k, blur, 0);
      for (long i = 0; i < N; ++i) {
        double b = blur[i] > eps ? blur[i] : eps;
        ratio[i] = (double)views[(long)v * N + i] / b;
      }
      rc |= conv_fft_d(ratio, s, kp, k, back, 1);
      for (long i = 0; i < N; ++i) acc[i] += back[i];
      kp += (long)k[0] * k[1] * k[2];
    }
    for (long i = 0; i < N; ++i) psi[i] *= acc[i] / V;
  }
  free(blur);
  free(ratio);
  free(back);
  free(acc);
  return rc;
}
