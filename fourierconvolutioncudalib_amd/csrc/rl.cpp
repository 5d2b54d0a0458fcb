// rl.cpp -- multiview Richardson-Lucy deconvolution on one device.
//
//   psi_0 = mean_v phi_v   (optional)
//   psi  <- psi * (1/V) sum_v corr( phi_v / max(conv(psi, P_v), eps), P_v )
//
// conv is the reference's convolution semantics (circular, kernel centred at
// floor(k/2); src/convolution3Dfft.cu:399-578); corr is its adjoint (the
// flipped PSF).  The reference has no deconvolution; this is the iterative
// caller of the hot path that the multiview workload runs.
//
// Fused form (every axis length has compile-time kernels): per iteration
//   x-R2C + y-FFT of psi                       2 passes over one volume
//   z fan-out: FFTz(psi) * FFTz(P_v) -> IFFTz  1 read, V spectra written
//   per view: inverse y-FFT                    V passes
//   x C2R -> ratio -> R2C (all views, 1 launch) V passes, no real volume
//   per view: y-FFT                            V passes
//   z sum over views of FFTz(r_v) conj(FFTz(P_v)) -> IFFTz   V reads, 1 write
//   inverse y-FFT, x-C2R with psi *= acc / V   2 passes
// Unfused form (any length): per view a convolution, the ratio, a
// correlation and an accumulation, then the update.
#include <algorithm>
#include <vector>

#include "../../include/fourierconvolution_amd.h"
#include "conv_driver.hpp"
#include "static_registry.hpp"

namespace fcamd {

void launch_rl_mean(float* psi, const float* const* d_views, int nviews, long n, hipStream_t st);
void launch_rl_ratio(float* ratio, const float* phi, const float* blur, float eps, long n,
                     hipStream_t st);
void launch_rl_accumulate(float* acc, const float* back, int first, long n, hipStream_t st);
void launch_rl_update(float* psi, const float* acc, float inv_views, long n, hipStream_t st);

namespace {

size_t a256(size_t b) { return (b + 255) / 256 * 256; }

struct RLProblem {
  ConvShape img;  // image shape (kernel fields unused)
  int V;
  std::vector<ConvShape> views;  // image shape + view kernel
};

RLProblem make_problem(int nviews, const int* imDim, const int* kernelDims) {
  FC_REQUIRE(nviews >= 1 && nviews <= kMaxViews, kInvalidArgument, "1..8 views supported");
  FC_REQUIRE(imDim && kernelDims, kInvalidArgument, "null dimension array");
  RLProblem p;
  p.V = nviews;
  p.img = ConvShape{imDim[0], imDim[1], imDim[2], 1, 1, 1, 0, 0, 0};
  for (int v = 0; v < nviews; ++v) {
    ConvShape s = p.img;
    s.kx = kernelDims[3 * v];
    s.ky = kernelDims[3 * v + 1];
    s.kz = kernelDims[3 * v + 2];
    s.cx = s.kx / 2;
    s.cy = s.ky / 2;
    s.cz = s.kz / 2;
    // shapes the drop-in call sends to the long-axis path take it in the
    // unfused form (the fused kernels exist for compile-time lengths only)
    if (needs_long_path(s))
      validate_long_shape(s);
    else
      validate_shape(s);
    p.views.push_back(s);
  }
  return p;
}

bool fused_available(const RLProblem& p) {
  StaticEntry* ex = static_lookup(p.img.nx);
  StaticEntry* ey = static_lookup(p.img.ny);
  StaticEntry* ez = static_lookup(p.img.nz);
  return ex && ey && ez && ex->rl.xratio && ez->rl.zfan && p.img.ny > 1;
}

// workspace: fused -> psi spectrum + V view spectra + V (Kx, Kxy);
// unfused -> conv workspace + blur + ratio + acc volumes + V (Kx, Kxy)
struct RLLayout {
  size_t spec, krows[kMaxViews], kplanes[kMaxViews], conv, vol;
  size_t total(const RLProblem& p, bool fused) const {
    size_t t = 0;
    for (int v = 0; v < p.V; ++v) t += krows[v] + kplanes[v];
    if (fused) return t + (size_t)(p.V + 1) * spec;
    return t + conv + 3 * vol;
  }
};

RLLayout rl_layout(const RLProblem& p) {
  RLLayout l{};
  const size_t nxp = padded_bins(p.img.nx);
  l.spec = a256(sizeof(float2) * nxp * p.img.ny * p.img.nz);
  l.conv = 0;
  for (int v = 0; v < p.V; ++v) {
    const ConvShape& s = p.views[v];
    l.krows[v] = a256(sizeof(float2) * nxp * s.ky * s.kz);
    l.kplanes[v] = a256(sizeof(float2) * nxp * s.ny * s.kz);
    l.conv = std::max(l.conv, needs_long_path(s) ? long_workspace_bytes(s)
                                                 : workspace_layout(s).total());
  }
  l.vol = a256(sizeof(float) * (size_t)p.img.voxels());
  return l;
}

void run_rl(const RLProblem& p, float* psi, const float* const* views, const float* const* kernels,
            int iterations, float eps, bool init, bool fused, void* ws, hipStream_t st) {
  const RLLayout l = rl_layout(p);
  const long n = (long)p.img.voxels();
  char* cur = static_cast<char*>(ws);
  auto take = [&](size_t b) {
    char* r = cur;
    cur += b;
    return r;
  };
  if (init) {
    // the view pointer table has to live in device memory for the kernel
    const float** d_views = reinterpret_cast<const float**>(take(a256(sizeof(float*) * kMaxViews)));
    FC_HIP_CHECK(hipMemcpyAsync(d_views, views, sizeof(float*) * p.V, hipMemcpyHostToDevice, st));
    launch_rl_mean(psi, d_views, p.V, n, st);
    FC_HIP_CHECK(hipStreamSynchronize(st));  // the host table goes out of scope
  } else {
    take(a256(sizeof(float*) * kMaxViews));
  }
  if (!fused) {
    void* cws = take(l.conv);
    float* blur = reinterpret_cast<float*>(take(l.vol));
    float* ratio = reinterpret_cast<float*>(take(l.vol));
    float* acc = reinterpret_cast<float*>(take(l.vol));
    auto conv = [&](const ConvShape& s, float* out, const float* in, const float* k, Op op) {
      if (needs_long_path(s))
        enqueue_convolution_long(s, out, in, k, op, cws, st);
      else
        enqueue_convolution(s, out, in, k, op, cws, st);
    };
    for (int it = 0; it < iterations; ++it) {
      for (int v = 0; v < p.V; ++v) {
        conv(p.views[v], blur, psi, kernels[v], Op::kConvolve);
        launch_rl_ratio(ratio, views[v], blur, eps, n, st);
        conv(p.views[v], blur, ratio, kernels[v], Op::kCorrelate);
        launch_rl_accumulate(acc, blur, v == 0 ? 1 : 0, n, st);
      }
      launch_rl_update(psi, acc, 1.f / (float)p.V, n, st);
    }
    FC_HIP_CHECK(hipGetLastError());
    return;
  }
  // fused iteration
  const ConvShape& s = p.img;
  StaticEntry* ex = static_lookup(s.nx);
  StaticEntry* ez = static_lookup(s.nz);
  float2* Spsi = reinterpret_cast<float2*>(take(l.spec));
  float2* Sv[kMaxViews];
  float2* Kx[kMaxViews];
  float2* Kxy[kMaxViews];
  for (int v = 0; v < p.V; ++v) Sv[v] = reinterpret_cast<float2*>(take(l.spec));
  for (int v = 0; v < p.V; ++v) {
    Kx[v] = reinterpret_cast<float2*>(take(l.krows[v]));
    Kxy[v] = reinterpret_cast<float2*>(take(l.kplanes[v]));
  }
  // kernel rows and planes of every view, once, through the merged
  // image + kernel passes (their image half transforms psi into Spsi, which
  // every iteration recomputes)
  for (int v = 0; v < p.V; ++v) {
    pass_x_forward(p.views[v], psi, Spsi, kernels[v], Kx[v], st);
    pass_y_forward(p.views[v], Spsi, Kx[v], Kxy[v], st);
  }
  const long nxp = padded_bins(s.nx);
  RLZArgs za{};
  za.fft = fft_desc(s.nz);
  za.nbins = half_bins(s.nx);
  za.s_row = nxp;
  za.s_plane = (long)s.ny * nxp;
  za.nviews = p.V;
  za.scale = (float)(1.0 / (double)s.voxels());
  za.src = Spsi;
  za.dst = Spsi;
  for (int v = 0; v < p.V; ++v) {
    za.view[v] = Sv[v];
    za.kplanes[v] = Kxy[v];
    za.kz[v] = p.views[v].kz;
    za.cz[v] = p.views[v].cz;
  }
  RLXArgs xa{};
  xa.fft = fft_desc(s.nx);
  xa.nrows = (long)s.ny * s.nz;
  xa.nbins = half_bins(s.nx);
  xa.srow = nxp;
  xa.rrow = s.nx;
  xa.eps = eps;
  xa.inv_views = 1.f / (float)p.V;
  for (int v = 0; v < p.V; ++v) {
    xa.spec[v] = Sv[v];
    xa.phi[v] = views[v];
  }
  xa.psi = psi;
  RLXArgs ua = xa;
  ua.spec[0] = Spsi;
  for (int it = 0; it < iterations; ++it) {
    pass_x_forward(s, psi, Spsi, nullptr, nullptr, st);
    pass_y_forward(s, Spsi, nullptr, nullptr, st);
    ez->rl.zfan(za, s.ny, st);
    for (int v = 0; v < p.V; ++v) pass_y_inverse(s, Sv[v], st);
    ex->rl.xratio(xa, p.V, st);
    for (int v = 0; v < p.V; ++v) pass_y_forward(s, Sv[v], nullptr, nullptr, st);
    ez->rl.zsum(za, s.ny, st);
    pass_y_inverse(s, Spsi, st);
    ex->rl.xupdate(ua, st);
  }
  FC_HIP_CHECK(hipGetLastError());
}

template <class F>
int rl_guarded(F&& f) {
  clear_error();
  try {
    f();
    return kOk;
  } catch (const Error& e) {
    set_error(e.code, e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error(kOutOfMemory, "host allocation failed");
    return kOutOfMemory;
  } catch (...) {
    set_error(kInternal, "unexpected internal error");
    return kInternal;
  }
}

}  // namespace
}  // namespace fcamd

using namespace fcamd;

extern "C" {

long long fcamd_rl_workspace_bytes(int nviews, const int* imDim, const int* kernelDims, int fused) {
  long long r = -1;
  (void)rl_guarded([&] {
    RLProblem p = make_problem(nviews, imDim, kernelDims);
    const bool f = fused != 0 && fused_available(p);
    r = (long long)(rl_layout(p).total(p, f) + a256(sizeof(float*) * kMaxViews));
  });
  return r;
}

int fcamd_rl_fused_available(int nviews, const int* imDim, const int* kernelDims) {
  int r = 0;
  (void)rl_guarded([&] { r = fused_available(make_problem(nviews, imDim, kernelDims)) ? 1 : 0; });
  return r;
}

int fcamd_rl_multiview_device(float* d_psi, const float* const* d_views, int nviews,
                              const int* imDim, const float* const* d_kernels,
                              const int* kernelDims, int iterations, float eps, int init_psi,
                              int fused, void* d_workspace, long long workspace_bytes,
                              void* stream) {
  return rl_guarded([&] {
    FC_REQUIRE(d_psi && d_views && d_kernels, kInvalidArgument, "null pointer");
    FC_REQUIRE(iterations >= 0, kInvalidArgument, "negative iteration count");
    FC_REQUIRE(eps > 0.f, kInvalidArgument, "eps must be positive");
    RLProblem p = make_problem(nviews, imDim, kernelDims);
    for (int v = 0; v < nviews; ++v)
      FC_REQUIRE(d_views[v] && d_kernels[v], kInvalidArgument, "null view or kernel");
    const bool f = fused != 0 && fused_available(p);
    const size_t need = rl_layout(p).total(p, f) + a256(sizeof(float*) * kMaxViews);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (d_workspace) {
      FC_REQUIRE(workspace_bytes >= (long long)need, kInvalidArgument, "workspace too small");
      run_rl(p, d_psi, d_views, d_kernels, iterations, eps, init_psi != 0, f, d_workspace, st);
    } else {
      int dev = 0;
      FC_HIP_CHECK(hipGetDevice(&dev));
      DevicePool& pool = device_pool(dev);
      std::lock_guard<std::mutex> lk(pool.mu);
      void* ws = pool_reserve(&pool.ws, &pool.ws_bytes, need);
      run_rl(p, d_psi, d_views, d_kernels, iterations, eps, init_psi != 0, f, ws, st);
      FC_HIP_CHECK(hipStreamSynchronize(st));
    }
  });
}

}  // extern "C"
