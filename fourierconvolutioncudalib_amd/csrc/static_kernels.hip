// static_kernels.hip -- dispatch to the compile-time-length kernels and
// their variants.
#include <atomic>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>

#include "../../include/fourierconvolution_amd.h"
#include "static_kernels.hpp"
#include "static_registry.hpp"

namespace fcamd {

namespace {
std::map<int, std::unique_ptr<StaticEntry>>& registry() {
  static std::map<int, std::unique_ptr<StaticEntry>> r;
  return r;
}
std::mutex& registry_mu() {
  static std::mutex m;
  return m;
}
// FCAMD_NO_STATIC=1 (or fcamd_use_static_kernels(0)) forces the
// runtime-length kernels
std::atomic<bool> g_disabled{false};
std::once_flag g_registered;
}  // namespace

// one per instantiation group of static_inst.hip (FC_GROUP 0..10)
void static_register_group_0();
void static_register_group_1();
void static_register_group_2();
void static_register_group_3();
void static_register_group_4();
void static_register_group_5();
void static_register_group_6();
void static_register_group_7();
void static_register_group_8();
void static_register_group_9();
void static_register_group_10();

namespace {
// Fills the registry on first use.  Calling the groups by name (instead of
// load-time constructors in their objects) keeps every group linked in when
// a program links the static archive.
void ensure_registered() {
  std::call_once(g_registered, [] {
    static_register_group_0();
    static_register_group_1();
    static_register_group_2();
    static_register_group_3();
    static_register_group_4();
    static_register_group_5();
    static_register_group_6();
    static_register_group_7();
    static_register_group_8();
    static_register_group_9();
    static_register_group_10();
  });
}
}  // namespace

void static_register(int L, const StaticEntry& e) {
  std::lock_guard<std::mutex> lk(registry_mu());
  registry()[L].reset(new StaticEntry(e));
}

StaticEntry* static_lookup(int L) {
  static const bool env_disabled = [] {
    const char* v = std::getenv("FCAMD_NO_STATIC");
    return v != nullptr && v[0] == '1';
  }();
  if (env_disabled || g_disabled) return nullptr;
  ensure_registered();
  std::lock_guard<std::mutex> lk(registry_mu());
  auto it = registry().find(L);
  return it == registry().end() ? nullptr : it->second.get();
}

int persistent_blocks(const void* kernel, int threads) {
  static std::mutex mu;
  static std::map<std::pair<int, const void*>, int> cache;
  int dev = 0;
  FC_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_pair(dev, kernel);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int per_cu = 0, cus = 0;
  FC_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0));
  FC_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int n = (per_cu < 1 ? 1 : per_cu) * (cus < 1 ? 1 : cus);
  cache[key] = n;
  return n;
}

namespace {
template <class V, class... A>
void run_selected(const std::vector<V>& vs, int sel, A&&... args) {
  if (sel > 0 && sel < (int)vs.size() && vs[sel].fn(args...)) return;
  vs[0].fn(args...);
}
}  // namespace

bool static_xfwd(const ConvShape& s, const FftDesc& fx, const float* src, float2* dst,
                 const float* kernel, float2* krows, hipStream_t st, const VPad* vp) {
  StaticEntry* e = static_lookup(s.nx);
  if (!e) return false;
  if (vp && vp->on && !e->xfwd_vpad) return false;
  XPass2 a{};
  a.m.fft = fx;
  a.m.nrows = src ? (long)s.ny * s.nz : 0;  // no image: the kernel's rows only
  a.m.nbins = half_bins(s.nx);
  a.m.src = src;
  a.m.dst = dst;
  a.m.src_row = s.nx;
  a.m.dst_row = padded_bins(s.nx);
  a.m.klen = s.nx;
  a.m.kcenter = 0;
  a.k = a.m;
  if (kernel) {
    a.k.nrows = (long)s.ky * s.kz;
    a.k.src = kernel;
    a.k.dst = krows;
    a.k.src_row = s.kx;
    a.k.klen = s.kx;
    a.k.kcenter = s.cx;
    a.k.kshift = 1;
  } else {
    a.k.nrows = 0;
  }
  if (vp && vp->on) {
    a.m.vp = *vp;
    return e->xfwd_vpad(a, st);
  }
  run_selected(e->xfwd, e->selected[kSxFwd], a, st);
  return true;
}

bool static_xinv(const ConvShape& s, const FftDesc& fx, const float2* src, float* dst,
                 hipStream_t st, const VPad* vp) {
  StaticEntry* e = static_lookup(s.nx);
  if (!e) return false;
  if (vp && vp->on && !e->xinv_crop) return false;
  XPassArgs a{};
  a.fft = fx;
  a.nrows = (long)s.ny * s.nz;
  a.nbins = half_bins(s.nx);
  a.csrc = src;
  a.rdst = dst;
  a.src_row = padded_bins(s.nx);
  a.dst_row = s.nx;
  if (vp && vp->on) {
    a.vp = *vp;
    return e->xinv_crop(a, st);
  }
  run_selected(e->xinv, e->selected[kSxInv], a, st);
  return true;
}

bool static_ycol(const ConvShape& s, const FftDesc& fy, float2* S, int dir, const float2* krows,
                 float2* kplanes, hipStream_t st) {
  StaticEntry* e = static_lookup(s.ny);
  if (!e) return false;
  const long nxp = padded_bins(s.nx);
  Col2 a{};
  a.m.fft = fy;
  a.m.nbins = half_bins(s.nx);
  a.m.src = S;
  a.m.dst = S;
  a.m.src_outer = a.m.dst_outer = (long)s.ny * nxp;
  a.m.src_elem = a.m.dst_elem = nxp;
  a.m.nvalid = s.ny;
  a.m.center = 0;
  a.m_outer = S ? s.nz : 0;  // no spectrum: the kernel's planes only
  a.k = a.m;
  a.k_outer = 0;
  if (krows) {
    a.k.src = krows;
    a.k.dst = kplanes;
    a.k.src_outer = (long)s.ky * nxp;
    a.k.src_elem = nxp;
    a.k.dst_outer = (long)s.ny * nxp;
    a.k.dst_elem = nxp;
    a.k.nvalid = s.ky;
    a.k.center = s.cy;
    a.k_outer = s.kz;
  }
  run_selected(e->col, e->selected[dir < 0 ? kSCol : kSColInv], a, dir, st);
  return true;
}

bool static_ycol_fill(const ConvShape& s, const FftDesc& fy, float2* S, int dir, int cus,
                      hipStream_t st) {
  StaticEntry* e = static_lookup(s.ny);
  if (!e) return false;
  const long nb = half_bins(s.nx);
  int best = -1, smallest = -1;
  for (int i = 0; i < (int)e->col.size(); ++i) {
    const int t = e->col[i].tile;
    if (t <= 0) continue;
    if (smallest < 0 || t < e->col[smallest].tile) smallest = i;
    if ((nb + t - 1) / t * s.nz >= cus && (best < 0 || t > e->col[best].tile)) best = i;
  }
  if (best < 0) best = smallest;
  if (best < 0) return static_ycol(s, fy, S, dir, nullptr, nullptr, st);
  const long nxp = padded_bins(s.nx);
  Col2 a{};
  a.m.fft = fy;
  a.m.nbins = nb;
  a.m.src = S;
  a.m.dst = S;
  a.m.src_outer = a.m.dst_outer = (long)s.ny * nxp;
  a.m.src_elem = a.m.dst_elem = nxp;
  a.m.nvalid = s.ny;
  a.m.center = 0;
  a.m_outer = s.nz;
  a.k = a.m;
  a.k_outer = 0;
  if (!e->col[best].fn(a, dir, st)) e->col[0].fn(a, dir, st);
  return true;
}

bool static_zfused(const ConvShape& s, const ZArgs& a, hipStream_t st) {
  StaticEntry* e = static_lookup(s.nz);
  if (!e) return false;
  run_selected(e->zfused, e->selected[kSZ], a, (long)s.ny, st);
  return true;
}

void static_kernels_allow_large_lds() {}

}  // namespace fcamd

using namespace fcamd;

extern "C" {

int fcamd_has_static_kernel(int pass, int length) {
  (void)pass;
  return static_lookup(length) != nullptr ? 1 : 0;
}

void fcamd_use_static_kernels(int enable) { g_disabled = enable == 0; }

int fcamd_static_lengths(int* lengths, int cap) {
  ensure_registered();
  std::lock_guard<std::mutex> lk(registry_mu());
  int n = 0;
  for (auto& kv : registry()) {
    if (n < cap && lengths) lengths[n] = kv.first;
    ++n;
  }
  return n;
}

int fcamd_static_variant_count(int pass, int length) {
  StaticEntry* e = static_lookup(length);
  return e ? e->count(pass) : 0;
}

const char* fcamd_static_variant_name(int pass, int length, int index) {
  StaticEntry* e = static_lookup(length);
  return e ? e->name(pass, index) : nullptr;
}

int fcamd_select_static_variant(int pass, int length, int index) {
  StaticEntry* e = static_lookup(length);
  if (!e || pass < 0 || pass >= kSNumPasses || index < 0 || index >= e->count(pass)) return kInvalidArgument;
  e->selected[pass] = index;
  return kOk;
}

int fcamd_selected_static_variant(int pass, int length) {
  StaticEntry* e = static_lookup(length);
  if (!e || pass < 0 || pass >= kSNumPasses) return -1;
  return e->selected[pass];
}

}  // extern "C"
