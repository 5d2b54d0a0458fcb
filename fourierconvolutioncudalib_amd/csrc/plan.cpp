// plan.cpp -- convolution plans: the kernel's spectrum computed once.
//
// A plan owns the workspace of one image shape and keeps the kernel's x-
// and y-transformed planes (Kxy) across calls; executing it runs only the
// image's passes, with the z pass taking the kernel column spectrum from
// the cached planes as usual.  This is the reuse the reference leaves as a
// TODO ("reuse the results", src/convolution3Dfft.cu:265,363).  The kernel's
// z spectrum is not stored: the fused z pass derives it per column tile,
// which moves fewer HBM bytes than reading an Nc-sized spectrum would.
// Shapes the drop-in call sends to the long-axis path (axes past one tile,
// large prime factors) keep the kernel's whole spectrum instead.
#include "../../include/fourierconvolution_amd.h"
#include "conv_driver.hpp"

namespace fcamd {
namespace {

struct ConvPlan {
  int dev = 0;
  ConvShape s{};
  Op op = Op::kConvolve;
  void* ws = nullptr;
  bool own = false;
  bool lng = false;  // long-axis path: the kernel's spectrum in the long layout
  float2 *S = nullptr, *Kx = nullptr, *Kxy = nullptr;
};

void plan_kernel(ConvPlan& p, const float* d_kernel, hipStream_t st) {
  FC_REQUIRE(d_kernel != nullptr, kInvalidArgument, "null kernel pointer");
  if (p.lng) {
    long_kernel_spectrum(p.s, d_kernel, p.ws, st);
    return;
  }
  // the kernel's rows and planes only (no image in these launches), by the
  // same kernels a plain call of this shape uses, so that executing the
  // plan reproduces the plain call bit for bit
  if (plane_path(p.s, -1)) {
    plane_forward(p.s, nullptr, nullptr, d_kernel, p.Kxy, st);
  } else {
    pass_x_forward(p.s, nullptr, nullptr, d_kernel, p.Kx, st);
    pass_y_forward(p.s, nullptr, p.Kx, p.Kxy, st);
  }
  FC_HIP_CHECK(hipGetLastError());
}

template <class F>
int plan_guarded(F&& f) {
  clear_error();
  try {
    f();
    return kOk;
  } catch (const Error& e) {
    set_error(e.code, e.msg);
    return e.code;
  } catch (...) {
    set_error(kInternal, "unexpected internal error");
    return kInternal;
  }
}

}  // namespace
}  // namespace fcamd

using namespace fcamd;

extern "C" {

void* fcamd_conv_plan_create(const int* imDim, const int* kernelDim, const float* d_kernel, int op,
                             void* d_workspace, long long workspace_bytes, void* stream) {
  ConvPlan* plan = nullptr;
  const int rc = plan_guarded([&] {
    FC_REQUIRE(imDim && kernelDim, kInvalidArgument, "null dimension array");
    FC_REQUIRE(op == FCAMD_CONVOLVE || op == FCAMD_CORRELATE, kInvalidArgument, "unknown op");
    ConvShape s{imDim[0], imDim[1], imDim[2], kernelDim[0], kernelDim[1], kernelDim[2],
                kernelDim[0] / 2, kernelDim[1] / 2, kernelDim[2] / 2};
    const bool lng = needs_long_path(s);
    if (lng)
      validate_long_shape(s);
    else
      validate_shape(s);
    driver_init();
    auto* p = new ConvPlan();
    p->s = s;
    p->lng = lng;
    p->op = op == FCAMD_CORRELATE ? Op::kCorrelate : Op::kConvolve;
    try {
      FC_HIP_CHECK(hipGetDevice(&p->dev));
      const size_t need = lng ? long_workspace_bytes(s) : workspace_layout(s).total();
      if (d_workspace) {
        FC_REQUIRE(workspace_bytes >= (long long)need, kInvalidArgument, "workspace too small");
        p->ws = d_workspace;
      } else {
        FC_HIP_CHECK(hipMalloc(&p->ws, need));
        p->own = true;
      }
      if (!lng) {
        const WorkspaceLayout wl = workspace_layout(s);
        char* base = static_cast<char*>(p->ws);
        p->S = reinterpret_cast<float2*>(base);
        p->Kx = reinterpret_cast<float2*>(base + wl.spectrum);
        p->Kxy = reinterpret_cast<float2*>(base + wl.spectrum + wl.krows);
        // twiddle tables now, so later executes allocate nothing (the long
        // path creates its own with the kernel's spectrum)
        int lens[5];
        for (int i = 0, n = table_lengths(s, lens); i < n; ++i) (void)fft_desc(lens[i]);
      }
      plan_kernel(*p, d_kernel, static_cast<hipStream_t>(stream));
    } catch (...) {
      if (p->own && p->ws) (void)hipFree(p->ws);
      delete p;
      throw;
    }
    plan = p;
  });
  return rc == kOk ? plan : nullptr;
}

int fcamd_conv_plan_set_kernel(void* plan, const float* d_kernel, void* stream) {
  return plan_guarded([&] {
    FC_REQUIRE(plan != nullptr, kInvalidArgument, "null plan");
    plan_kernel(*static_cast<ConvPlan*>(plan), d_kernel, static_cast<hipStream_t>(stream));
  });
}

int fcamd_conv_plan_execute(void* plan, float* d_out, const float* d_image, void* stream) {
  return plan_guarded([&] {
    FC_REQUIRE(plan != nullptr, kInvalidArgument, "null plan");
    FC_REQUIRE(d_out && d_image, kInvalidArgument, "null device pointer");
    ConvPlan& p = *static_cast<ConvPlan*>(plan);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (p.lng) {
      long_apply_spectrum(p.s, d_out, d_image, p.op, p.ws, st);
      return;
    }
    if (plane_path(p.s, -1))
      plane_forward(p.s, d_image, p.S, nullptr, nullptr, st);
    else
      forward_xy(p.s, d_image, p.S, nullptr, nullptr, nullptr, st);
    z_pass(p.s, p.S, p.Kxy, false, p.op == Op::kCorrelate, st);
    if (plane_path(p.s, +1))
      plane_inverse(p.s, p.S, d_out, st);
    else
      inverse_yx(p.s, p.S, d_out, st);
    FC_HIP_CHECK(hipGetLastError());
  });
}

void fcamd_conv_plan_destroy(void* plan) {
  auto* p = static_cast<ConvPlan*>(plan);
  if (!p) return;
  if (p->own && p->ws) (void)hipFree(p->ws);
  delete p;
}

}  // extern "C"
