/* fc_oracle.h -- CPU oracle for the 3D FFT convolution hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py use it, and only as the checker.
 *
 * Every function restates a piece of the reference (citations are relative to
 * the reference repository root) in plain C.  Volumes are C-order
 * [nz][ny][nx] float/double arrays, x fastest; shapes are passed as C-order
 * {nz, ny, nx} unless the name says otherwise.
 */
#ifndef FC_ORACLE_H
#define FC_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* tests/test_algorithms.hpp:9-58 -- direct (linear, flipped-kernel)
 * convolution over the interior [off, shape-off) of `image`, float
 * accumulation in z->y->x / kz->ky->kx loop order, so results are
 * bit-identical to the reference template.  Cells outside the interior of
 * `result` are left untouched. */
void fco_convolve_host(const float* image, const int* image_shape,
                       const float* kernel, const int* kernel_shape,
                       float* result, const int* offsets);

/* src/convolution3Dfft.cu:128-166 (fftShiftKernel) with the axis roles the
 * rest of convolution3DfftCUDAInPlace uses (x fastest): kernel tap
 * (z,y,x) lands at ((z-kz/2) mod nz, (y-ky/2) mod ny, (x-kx/2) mod nx) of a
 * zeroed image-sized volume.  Returns -1 if a kernel extent exceeds the
 * image extent (the reference writes out of bounds there). */
int fco_fftshift_pad(const float* kernel, const int* kernel_shape,
                     double* padded, const int* image_shape);

/* Semantics of convolution3DfftCUDAInPlace (src/convolution3Dfft.cu:399-578,
 * multiply :41-62 without conjugation, scale 1/N at :531) evaluated
 * directly in FP64: out[i] = sum_j k[j] * im[(i - (j - floor(k/2))) mod n].
 * O(N*K); for small shapes. */
int fco_circular_conv_direct(const float* image, const int* image_shape,
                             const float* kernel, const int* kernel_shape,
                             double* out);

/* In-place FP64 complex DFT of any length n (mixed radix, prime factors by
 * direct summation); data is interleaved re/im, element i at data[2*i*stride].
 * sign = -1 forward (cuFFT's forward convention), +1 inverse; unnormalised. */
/* Linear convolution (correlate = 0) or correlation (1) of an image
 * extended beyond its edges by boundary rule `mode` (0 periodic, 1 zero,
 * 2 nearest, 3 reflect, 4 mirror); kernel centred at k/2, any size. */
int fco_boundary_conv_direct(const float* image, const int* image_shape, const float* kernel,
                             const int* kernel_shape, int mode, int correlate, double* out);

void fco_fft1d(double* data, int n, long stride, int sign);

/* FP64 3D complex DFT in place over a C-order [nz][ny][nx] interleaved
 * complex volume; sign as above, unnormalised. */
void fco_fft3d(double* data, const int* shape, int sign);

/* FP64 FFT circular convolution with the same semantics as
 * fco_circular_conv_direct: IFFT(FFT(im) * FFT(fftshift_pad(k))) / N.
 * conj_kernel != 0 gives the adjoint (correlation with the kernel). */
int fco_fft_conv(const float* image, const int* image_shape,
                 const float* kernel, const int* kernel_shape, double* out,
                 int conj_kernel);

/* Adjoint with respect to the kernel of out = conv(im, k) at the loss
 * gradient g (both image-sized):
 *   dk[j] = sum_i g[i] * im[(i - (j - floor(k/2))) mod n]   (FP64, direct). */
int fco_conv_grad_kernel(const double* g, const float* image,
                         const int* image_shape, const int* kernel_shape,
                         double* dk);

/* tests/padd_utils.h:100-115 (zero_padd extents/offsets) and :157-171
 * (insert_at_offsets): extents = shape + 2*factor*(k/2), offsets =
 * factor*(k/2); the image is copied into a zeroed padded volume. */
void fco_zero_padd_extents(const int* image_shape, const int* kernel_shape,
                           int factor, int* extents, int* offsets);
void fco_insert_at_offsets(const float* image, const int* image_shape,
                           float* padded, const int* extents,
                           const int* offsets);

/* tests/test_utils.hpp:76-88: sqrt(sum (a-b)^2) / N in double. */
double fco_l2norm_ref(const float* a, const float* b, long n);

/* FP64 multiview Richardson-Lucy with V views sharing one image shape:
 *   psi <- psi * (1/V) sum_v conv_adj(phi_v / conv(psi, P_v), P_v)
 * with psi0 = mean of views; `eps` guards the division.  Kernels are
 * concatenated, kernel_shapes holds 3 ints per view. */
int fco_rl_multiview(const float* views, int nviews, const int* image_shape,
                     const float* kernels, const int* kernel_shapes,
                     int iterations, double eps, double* psi);

#ifdef __cplusplus
}
#endif

#endif /* FC_ORACLE_H */
