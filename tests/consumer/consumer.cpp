// consumer.cpp -- a C++ caller of libFourierConvolutionCUDALib: the same
// calls as consumer.c through the header's C++ declarations (extern "C"
// via FUNCTION_PREFIX), releasing the legacy result with delete[] as the
// reference's callers must (src/convolution3Dfft.cu:240,327).
// Usage: consumer_cxx IMAGE NX NY NZ KERNEL KX KY KZ OUT_INPLACE OUT_LEGACY
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "consumer_io.h"
#include "convolution3Dfft.h"
#include "fourierconvolution_amd.h"

int main(int argc, char** argv) {
  if (argc != 11) {
    std::fprintf(stderr, "usage: %s IMAGE NX NY NZ KERNEL KX KY KZ OUT_INPLACE OUT_LEGACY\n",
                 argv[0]);
    return 2;
  }
  int imDim[3], kDim[3], revIm[3], revK[3];
  for (int d = 0; d < 3; ++d) {
    imDim[d] = std::atoi(argv[2 + d]);
    kDim[d] = std::atoi(argv[6 + d]);
  }
  for (int d = 0; d < 3; ++d) {
    revIm[d] = imDim[2 - d];
    revK[d] = kDim[2 - d];
  }
  const long n = (long)imDim[0] * imDim[1] * imDim[2];
  const long nk = (long)kDim[0] * kDim[1] * kDim[2];
  float* im = read_floats(argv[1], n);
  float* kernel = read_floats(argv[5], nk);
  if (!im || !kernel) return 2;
  std::printf("header_constants_ok=%d\n", header_constants_ok());
  const int dev = selectDeviceWithHighestComputeCapability();
  std::printf("device=%d\n", dev);
  if (dev < 0) return 3;
  std::vector<char> name(256, 'x');
  getNameDeviceCUDA(dev, name.data());
  name[255] = 0;
  std::printf("name=%s\n", name.data());
  std::printf("cc=%d.%d\n", getCUDAcomputeCapabilityMajorVersion(dev),
              getCUDAcomputeCapabilityMinorVersion(dev));
  imageType* legacy = convolution3DfftCUDA(im, revIm, kernel, revK, dev);
  if (!legacy) {
    std::fprintf(stderr, "convolution3DfftCUDA: %s\n", fcamd_last_error_message());
    return 4;
  }
  convolution3DfftCUDAInPlace(im, imDim, kernel, kDim, dev);
  if (fcamd_last_error() != 0) {
    std::fprintf(stderr, "convolution3DfftCUDAInPlace: %s\n", fcamd_last_error_message());
    return 4;
  }
  if (write_floats(argv[9], im, n) || write_floats(argv[10], legacy, n)) return 5;
  delete[] legacy;
  std::free(im);
  std::free(kernel);
  std::printf("status=ok\n");
  return 0;
}
