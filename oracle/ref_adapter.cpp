// Test infrastructure only: builds the reference's own host convolution
// template (tests/test_algorithms.hpp:9-58, `fourierconvolution::convolve`)
// from the reference source tree into oracle/_ref/libref_convolve.so.
//
// The template only needs an image type with num_elements(), shape() and
// [z][y][x] access.  `tests/test_algorithms.hpp:5` includes
// `image_stack_utils.h`, which holds Boost typedefs and printers that
// `convolve` never uses; the recipe in oracle/Makefile predefines that
// header's include guard so the Boost-free template compiles alone.  No
// reference source is copied: this file supplies the caller-side array type
// and a C entry point around the unchanged template.
#include <algorithm>  // std::min_element, normally pulled in through Boost
#include <cmath>
#include <cstddef>
#include <iostream>
#include <vector>

#include "test_algorithms.hpp"  // resolved via -I<reference>/tests

namespace {

// A C-order [nz][ny][nx] float volume viewed through pointer + shape, exactly
// the access pattern the template uses (boost::multi_array in the reference).
struct Plane1 {
  float* p;
  float& operator[](long i) const { return p[i]; }
};
struct Plane2 {
  float* p;
  std::size_t nx;
  Plane1 operator[](long i) const { return Plane1{p + i * (long)nx}; }
};
struct Volume {
  float* data_;
  std::size_t shape_[3];  // {nz, ny, nx}
  std::size_t num_elements() const { return shape_[0] * shape_[1] * shape_[2]; }
  const std::size_t* shape() const { return shape_; }
  Plane2 operator[](long i) const {
    return Plane2{data_ + i * (long)(shape_[1] * shape_[2]), shape_[2]};
  }
};

}  // namespace

extern "C" {

// result must be a buffer of the image's shape; only the interior
// [off, shape-off) per axis is written (as in the reference template).
// offsets are {oz, oy, ox}, shapes are C-order {nz, ny, nx}.
void ref_convolve(float* image, const int* image_shape, float* kernel,
                  const int* kernel_shape, float* result, const int* offsets) {
  Volume im{image, {(std::size_t)image_shape[0], (std::size_t)image_shape[1],
                    (std::size_t)image_shape[2]}};
  Volume k{kernel, {(std::size_t)kernel_shape[0], (std::size_t)kernel_shape[1],
                    (std::size_t)kernel_shape[2]}};
  Volume out{result, {(std::size_t)image_shape[0], (std::size_t)image_shape[1],
                      (std::size_t)image_shape[2]}};
  std::vector<int> off(offsets, offsets + 3);
  fourierconvolution::convolve(im, k, out, off);
}

}  // extern "C"
