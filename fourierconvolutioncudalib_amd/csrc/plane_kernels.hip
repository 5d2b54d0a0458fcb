// plane_kernels.hip -- instantiations and dispatch of the whole-plane x/y
// kernels (plane_kernels.hpp); the wide planes are instantiated in
// plane_kernels_wide.hip.
#include <atomic>
#include <cstdlib>

#include "conv_driver.hpp"
#include "plane_kernels.hpp"
#include "plane_registry.hpp"
#include "static_registry.hpp"

namespace fcamd {
namespace {

template <int NX, int NY>
constexpr PlaneFn split_fwd() {
  if constexpr (plane_fwd2_ok<NX, NY>())
    return &plane_launch_fwd2<NX, NY>;
  else
    return nullptr;
}
template <int NX, int NY>
constexpr PlaneFn split_inv() {
  if constexpr (plane_inv2_ok<NX, NY>())
    return &plane_launch_inv2<NX, NY>;
  else
    return nullptr;
}

#define FC_PLANE(NX, NY)                                                                 \
  {NX, NY, &plane_launch_fwd<NX, NY>, &plane_launch_inv<NX, NY>, split_fwd<NX, NY>(), \
   split_inv<NX, NY>(), false, nullptr, nullptr}
#define FC_PLANE_ROW(NX)                                                          \
  FC_PLANE(NX, 64), FC_PLANE(NX, 80), FC_PLANE(NX, 96), FC_PLANE(NX, 112),        \
      FC_PLANE(NX, 120), FC_PLANE(NX, 128)
// Plane extents (nx, ny) with whole-plane kernels: pairs of compile-time
// lengths (nx a multiple of 4) whose row pairs all fit one exchange group
// next to the half-spectrum plane in LDS (and whose columns fit in at most
// two groups).  Planes that need two row groups measured mixed (128 x 192:
// 48.6 -> 52.8 us; 192 x 128: 64.5 -> 58.5 us) or slower (160 x 160 at 256
// planes: 78.8 -> 90.0 us); single-group planes faster (64 x 256: 62.7 ->
// 40.6 us, 160 x 96: 56.2 -> 43.8 us), so only those are instantiated.
const PlaneEntry kPlanes[] = {
    FC_PLANE_ROW(64), FC_PLANE_ROW(80), FC_PLANE_ROW(96), FC_PLANE_ROW(112), FC_PLANE_ROW(120),
    FC_PLANE_ROW(128),
    FC_PLANE(64, 144), FC_PLANE(64, 160), FC_PLANE(64, 192), FC_PLANE(64, 224),
    FC_PLANE(64, 240), FC_PLANE(64, 256), FC_PLANE(80, 144), FC_PLANE(80, 160),
    FC_PLANE(80, 192), FC_PLANE(80, 224), FC_PLANE(96, 144), FC_PLANE(96, 160),
    FC_PLANE(112, 144), FC_PLANE(112, 160), FC_PLANE(120, 144), FC_PLANE(128, 144),
    FC_PLANE(144, 64), FC_PLANE(144, 80), FC_PLANE(144, 96), FC_PLANE(144, 112),
    FC_PLANE(144, 120), FC_PLANE(144, 128), FC_PLANE(160, 64), FC_PLANE(160, 80),
    FC_PLANE(160, 96), FC_PLANE(160, 112), FC_PLANE(192, 64), FC_PLANE(192, 80),
    FC_PLANE(192, 96), FC_PLANE(224, 64), FC_PLANE(224, 80), FC_PLANE(240, 64),
    FC_PLANE(256, 64),
};
#undef FC_PLANE_ROW
#undef FC_PLANE

// FCAMD_NO_PLANES=1 keeps the separate x and y passes
bool planes_enabled() {
  static const bool on = [] {
    const char* v = std::getenv("FCAMD_NO_PLANES");
    return !(v && v[0] && v[0] != '0');
  }();
  return on;
}

const PlaneEntry* plane_entry(const ConvShape& s) {
  if (!planes_enabled()) return nullptr;
  for (const PlaneEntry& e : kPlanes)
    if (e.nx == s.nx && e.ny == s.ny) return &e;
  return wide_plane_entry(s.nx, s.ny);
}

// FCAMD_PLANE_SPLIT: 0 = never; 1 = always (where the half lengths have
// plans); f / i = the forward / the inverse only; unset or anything else =
// automatic
int split_mode() {
  static const int mode = [] {
    const char* v = std::getenv("FCAMD_PLANE_SPLIT");
    if (!v || !v[0] || v[1]) return -1;
    switch (v[0]) {
      case '0': return 0;
      case 'f': return 1;
      case 'i': return 2;
      case '1': return 3;
      default: return -1;
    }
  }();
  return mode;
}

}  // namespace

int cu_count() {
  constexpr int kDevices = 64;
  static std::atomic<int> cache[kDevices];
  int dev = 0;
  FC_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kDevices) dev = 0;
  int n = cache[dev].load(std::memory_order_relaxed);
  if (n == 0) {
    FC_HIP_CHECK(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    cache[dev].store(n, std::memory_order_relaxed);
  }
  return n;
}

namespace {

// The split halves each workgroup's work but doubles the workgroups: it pays
// while both parts of every plane run in one wave (measured on MI355X, 256
// CUs, GPU time per convolution: 128 x 128 x 64 28.58 -> 25.80 us with both
// directions split; 128 x 128 x 128, 143 forward planes: forward split 36.91
// -> 41.48 us, inverse split 36.91 -> 35.50 us; 128 x 128 x 256: inverse
// split 52.07 -> 56.04 us).  The forward kernel holds one workgroup per CU
// (its LDS), the inverse two, but the inverse's second workgroup per CU
// measured slower, so both directions take one wave of one per CU.
bool split_auto(const ConvShape& s, int dir) {
  const int planes = dir < 0 ? s.nz + s.kz : s.nz;
  return 2 * planes <= cu_count();
}

// Wide planes replace two passes through HBM rather than one workgroup per
// plane.  Measured on MI355X (GPU time per convolution, separate passes
// against the wide kernel in one direction): the forward gains while its
// workgroups fit one round (256 x 256 x 64, 15^3: 64.6 -> 61.3 us) and
// loses at two (256 x 256 x 128, 21^3, 298 workgroups of one per CU: 86.9
// -> 96.1 us); the inverse gains at one round (86.9 -> 77.7 us) and at two
// (256 x 256 x 256: 133.0 -> 124.9 us).
bool split_auto_wide(const PlaneEntry& e, const ConvShape& s, int dir) {
  if (dir < 0) return 2 * (s.nz + s.kz) <= cu_count() * e.occ_fwd2();
  return 2 * s.nz <= 2 * cu_count() * e.occ_inv2();
}

}  // namespace

bool plane_split(const ConvShape& s, int dir) {
  const PlaneEntry* e = plane_entry(s);
  if (e == nullptr || (dir < 0 ? e->fwd2 : e->inv2) == nullptr) return false;
  // the per-stage tables of the compile-time plans carry the twiddles
  if (!static_lookup(s.nx) || !static_lookup(s.ny)) return false;
  switch (split_mode()) {
    case 0: return false;
    case 1: return dir < 0;
    case 2: return dir > 0;
    case 3: return true;
    default: return e->wide ? split_auto_wide(*e, s, dir) : split_auto(s, dir);
  }
}

bool plane_path(const ConvShape& s, int dir) {
  const PlaneEntry* e = plane_entry(s);
  if (e == nullptr || !static_lookup(s.nx) || !static_lookup(s.ny)) return false;
  return (dir < 0 ? e->fwd : e->inv) != nullptr || plane_split(s, dir);
}

int plane_half_length(const ConvShape& s, int dir) { return dir < 0 ? s.nx / 2 : s.ny / 2; }

void plane_launch(const ConvShape& s, const PlaneArgs& a, int dir, hipStream_t st) {
  const PlaneEntry* e = plane_entry(s);
  FC_REQUIRE(plane_path(s, dir), kInternal, "no whole-plane kernels for this plane extent");
  if (plane_split(s, dir)) {
    FC_REQUIRE((dir < 0 ? a.hx_stw : a.hy_stw) != nullptr, kInternal,
               "split plane kernels without the half-length table");
    (dir < 0 ? e->fwd2 : e->inv2)(a, st);
  } else {
    (dir < 0 ? e->fwd : e->inv)(a, st);
  }
}

}  // namespace fcamd
