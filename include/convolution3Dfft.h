/* convolution3Dfft.h -- public C ABI of libFourierConvolutionCUDALib, MI355X
 * (gfx950) build.  Every declaration here keeps the name, argument types and
 * return type of the reference header so existing callers (Fiji via JNA, C++
 * test programs) link against this library unchanged.  Each entry cites the
 * reference declaration it replaces (src/convolution3Dfft.h).
 *
 * Layout contract (src/convolution3Dfft.cu:394-396):
 *   volumes are C-order float arrays and dims are given as {x, y, z} with
 *   imDim[0] the contiguous (fastest) axis, for convolution3DfftCUDAInPlace.
 *   The two legacy functions that return a new buffer take dims with
 *   imDim[2] the contiguous axis (pos = z + imDim[2]*(y + imDim[1]*x)).
 *
 * Semantics: circular 3D convolution of the image with the kernel
 * zero-padded to the image size and centred at floor(k/2) on every axis,
 * i.e. out[i] = sum_j k[j] * im[(i - (j - floor(k/2))) mod n]; the kernel is
 * not flipped.
 *
 * Errors never terminate the process: a failing call leaves `im` untouched
 * (the returning variants return NULL) and records a code and message that
 * fcamd_last_error() / fcamd_last_error_message() report
 * (fourierconvolution_amd.h).
 */
#ifndef __CONVOLUTION_3D_FFT_H__
#define __CONVOLUTION_3D_FFT_H__

#include "FourierConvolutionCUDALib_Export.h"

#ifdef __cplusplus
#define FUNCTION_PREFIX extern "C" FourierConvolutionCUDALib_EXPORT
#else
#define FUNCTION_PREFIX extern FourierConvolutionCUDALib_EXPORT
#endif

typedef float imageType; /* src/convolution3Dfft.h:17 */
static const int MAX_THREADS_CUDA = 1024; /* :18, kept for source compatibility */
static const int MAX_BLOCKS_CUDA = 65535; /* :19 */
static const int dimsImage = 3;           /* :21 */

/* src/convolution3Dfft.h:29-32.  `kernel` is image-sized and already in
 * wrap-around order (tap 0 at index 0); dims have imDim[2] contiguous.
 * Returns a new[]-allocated float[N] the caller releases with delete[]
 * (or fcamd_free_result); NULL on error. */
FUNCTION_PREFIX imageType* convolution3DfftCUDA_test(imageType* im, int* imDim,
                                                     imageType* kernel, int devCUDA);

/* src/convolution3Dfft.h:41-45.  Same operation as the in-place variant with
 * dims given z-fastest (imDim[2], kernelDim[2] contiguous); `im` is not
 * modified; returns new float[N] (delete[] / fcamd_free_result); NULL on
 * error. */
FUNCTION_PREFIX imageType* convolution3DfftCUDA(imageType* im, int* imDim, imageType* kernel,
                                                int* kernelDim, int devCUDA);

/* src/convolution3Dfft.h:56 -- the hot path.  Host pointers; `im`
 * (imDim = {nx, ny, nz}, x contiguous) is overwritten with the convolution,
 * `kernel` (kernelDim = {kx, ky, kz}) is read only.  Kernel extents must not
 * exceed image extents. */
FUNCTION_PREFIX void convolution3DfftCUDAInPlace(imageType* im, int* imDim, imageType* kernel,
                                                 int* kernelDim, int devCUDA);

/* src/convolution3Dfft.h:66-72 and src/standardCUDAfunctions.cu:13-71.
 * "Compute capability" is the gfx major/minor the HIP runtime reports
 * (9 / 5 on MI355X).  Failing queries return -1 (0 for the device count). */
FUNCTION_PREFIX int selectDeviceWithHighestComputeCapability();
FUNCTION_PREFIX int getCUDAcomputeCapabilityMinorVersion(int devCUDA);
FUNCTION_PREFIX int getCUDAcomputeCapabilityMajorVersion(int devCUDA);
FUNCTION_PREFIX int getNumDevicesCUDA();
/* writes 256 bytes (the device name, NUL padded) into `name` */
FUNCTION_PREFIX void getNameDeviceCUDA(int devCUDA, char* name);
FUNCTION_PREFIX long long int getMemDeviceCUDA(int devCUDA);

/* src/convolution3Dfft.h:73 (src/convolution3Dfft.cu:581-584): the runtime
 * version the library was built against; here HIP_VERSION
 * (major*10000000 + minor*100000 + patch). */
FUNCTION_PREFIX int cuda_version();

/* src/convolution3Dfft.h:75-86 (src/convolution3Dfft.cu:587-603, there
 * cuFFT's work area): the device memory in MiB (rounded up) a host-pointer
 * convolution of an image of `shape` (len 1..3, last entry contiguous)
 * allocates on the current device for a one-voxel kernel, within the device
 * budget (fcamd_host_device_bytes in fourierconvolution_amd.h gives the
 * figure for any kernel); -1 for an invalid request. */
FUNCTION_PREFIX int gpu_mem_needed_mb(int* shape, int len);

#endif /* __CONVOLUTION_3D_FFT_H__ */
