// static_kernels.hip -- dispatch to the compile-time-length kernels.
#include <atomic>
#include <cstdlib>
#include <map>
#include <mutex>

#include "../../include/fourierconvolution_amd.h"
#include "static_kernels.hpp"
#include "static_registry.hpp"

namespace fcamd {

namespace {
}  // namespace
std::map<int, StaticEntry>& registry() {
  static std::map<int, StaticEntry> r;
  return r;
}
std::mutex& registry_mu() {
  static std::mutex m;
  return m;
}
std::atomic<bool> g_disabled{false};  // FCAMD_NO_STATIC=1 also forces the runtime-length kernels

void static_register(int L, const StaticEntry& e) {
  std::lock_guard<std::mutex> lk(registry_mu());
  registry()[L] = e;
}

const StaticEntry* static_lookup(int L) {
  static const bool disabled = [] {
    const char* v = std::getenv("FCAMD_NO_STATIC");
    return v != nullptr && v[0] == '1';
  }();
  if (disabled || g_disabled) return nullptr;
  std::lock_guard<std::mutex> lk(registry_mu());
  auto it = registry().find(L);
  return it == registry().end() ? nullptr : &it->second;
}

bool static_xfwd(const ConvShape& s, const FftDesc& fx, const float* src, float2* dst,
                 hipStream_t st) {
  const StaticEntry* e = static_lookup(s.nx);
  if (!e) return false;
  XPassArgs a{};
  a.fft = fx;
  a.nrows = (long)s.ny * s.nz;
  a.nbins = half_bins(s.nx);
  a.src = src;
  a.dst = dst;
  a.src_row = s.nx;
  a.dst_row = padded_bins(s.nx);
  e->xfwd(a, (a.nrows + 1) / 2, st);
  return true;
}

bool static_xinv(const ConvShape& s, const FftDesc& fx, const float2* src, float* dst,
                 hipStream_t st) {
  const StaticEntry* e = static_lookup(s.nx);
  if (!e) return false;
  XPassArgs a{};
  a.fft = fx;
  a.nrows = (long)s.ny * s.nz;
  a.nbins = half_bins(s.nx);
  a.csrc = src;
  a.rdst = dst;
  a.src_row = padded_bins(s.nx);
  a.dst_row = s.nx;
  e->xinv(a, (a.nrows + 1) / 2, st);
  return true;
}

bool static_ycol(const ConvShape& s, const FftDesc& fy, float2* S, int dir, hipStream_t st) {
  const StaticEntry* e = static_lookup(s.ny);
  if (!e) return false;
  const long nxp = padded_bins(s.nx);
  ColArgs a{};
  a.fft = fy;
  a.nbins = half_bins(s.nx);
  a.src = S;
  a.dst = S;
  a.src_outer = a.dst_outer = (long)s.ny * nxp;
  a.src_elem = a.dst_elem = nxp;
  a.nvalid = s.ny;
  a.center = 0;
  e->col(a, dir, s.nz, st);
  return true;
}

bool static_zfused(const ConvShape& s, const ZArgs& a, hipStream_t st) {
  const StaticEntry* e = static_lookup(s.nz);
  if (!e) return false;
  e->zfused(a, s.ny, st);
  return true;
}

void static_kernels_allow_large_lds() {}

}  // namespace fcamd

extern "C" int fcamd_has_static_kernel(int pass, int length) {
  (void)pass;
  return fcamd::static_lookup(length) != nullptr ? 1 : 0;
}

extern "C" void fcamd_use_static_kernels(int enable) { fcamd::g_disabled = enable == 0; }

extern "C" int fcamd_static_lengths(int* lengths, int cap) {
  std::lock_guard<std::mutex> lk(fcamd::registry_mu());
  int n = 0;
  for (auto& kv : fcamd::registry()) {
    if (n < cap) lengths[n] = kv.first;
    ++n;
  }
  return n;
}
