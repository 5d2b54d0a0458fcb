/* consumer.c -- a C99 caller of libFourierConvolutionCUDALib, written the way
 * the reference's callers use it (tests/test_gpu_convolve.cpp:233-234 pass
 * reversed C-order shapes and select the device first): it includes the
 * public headers, links the library (shared or static) and calls
 *   convolution3DfftCUDAInPlace   on the padded image, dims {x, y, z}
 *   convolution3DfftCUDA          on the same image, dims z-fastest
 *   the device queries
 * Usage: consumer IMAGE NX NY NZ KERNEL KX KY KZ OUT_INPLACE OUT_LEGACY
 * Prints key=value lines; exit status 0 when every call succeeded. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "consumer_io.h"
#include "convolution3Dfft.h"
#include "fourierconvolution_amd.h"

int main(int argc, char** argv) {
  int imDim[3], kDim[3], revIm[3], revK[3], d, dev, len;
  long n, nk;
  float *im, *kernel, *legacy;
  char name[256];
  int shape[3];
  if (argc != 11) {
    fprintf(stderr, "usage: %s IMAGE NX NY NZ KERNEL KX KY KZ OUT_INPLACE OUT_LEGACY\n", argv[0]);
    return 2;
  }
  for (d = 0; d < 3; ++d) {
    imDim[d] = atoi(argv[2 + d]);
    kDim[d] = atoi(argv[6 + d]);
  }
  for (d = 0; d < 3; ++d) {
    revIm[d] = imDim[2 - d];
    revK[d] = kDim[2 - d];
  }
  n = (long)imDim[0] * imDim[1] * imDim[2];
  nk = (long)kDim[0] * kDim[1] * kDim[2];
  im = read_floats(argv[1], n);
  kernel = read_floats(argv[5], nk);
  if (!im || !kernel) {
    fprintf(stderr, "cannot read inputs\n");
    return 2;
  }
  printf("header_constants_ok=%d\n", header_constants_ok());
  printf("num_devices=%d\n", getNumDevicesCUDA());
  dev = selectDeviceWithHighestComputeCapability();
  printf("device=%d\n", dev);
  if (dev < 0) return 3;
  memset(name, 'x', sizeof(name));
  getNameDeviceCUDA(dev, name);
  name[255] = 0;
  printf("name=%s\n", name);
  printf("cc=%d.%d\n", getCUDAcomputeCapabilityMajorVersion(dev),
         getCUDAcomputeCapabilityMinorVersion(dev));
  printf("mem=%lld\n", getMemDeviceCUDA(dev));
  printf("version=%d\n", cuda_version());
  shape[0] = imDim[2];
  shape[1] = imDim[1];
  shape[2] = imDim[0];
  len = 3;
  printf("mem_needed_mb=%d\n", gpu_mem_needed_mb(shape, len));

  /* z-fastest legacy entry point first: it must leave `im` untouched */
  legacy = convolution3DfftCUDA(im, revIm, kernel, revK, dev);
  if (!legacy) {
    fprintf(stderr, "convolution3DfftCUDA: %s\n", fcamd_last_error_message());
    return 4;
  }
  convolution3DfftCUDAInPlace(im, imDim, kernel, kDim, dev);
  if (fcamd_last_error() != 0) {
    fprintf(stderr, "convolution3DfftCUDAInPlace: %s\n", fcamd_last_error_message());
    return 4;
  }
  if (write_floats(argv[9], im, n) || write_floats(argv[10], legacy, n)) return 5;
  fcamd_free_result(legacy); /* C cannot call delete[] */
  free(im);
  free(kernel);
  printf("status=ok\n");
  return 0;
}
