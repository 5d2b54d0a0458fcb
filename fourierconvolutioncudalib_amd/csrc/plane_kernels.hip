// plane_kernels.hip -- instantiations and dispatch of the whole-plane x/y
// kernels (plane_kernels.hpp).
#include <cstdlib>

#include "conv_driver.hpp"
#include "plane_kernels.hpp"
#include "static_registry.hpp"

namespace fcamd {
namespace {

using PlaneFn = void (*)(const PlaneArgs&, hipStream_t);
struct PlaneEntry {
  int nx, ny;
  PlaneFn fwd, inv;
};

#define FC_PLANE(NX, NY) {NX, NY, &plane_launch_fwd<NX, NY>, &plane_launch_inv<NX, NY>}
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
  return nullptr;
}

}  // namespace

bool plane_path_available(const ConvShape& s) {
  // the per-stage tables of the compile-time plans carry the twiddles
  return plane_entry(s) != nullptr && static_lookup(s.nx) && static_lookup(s.ny);
}

void plane_launch(const ConvShape& s, const PlaneArgs& a, int dir, hipStream_t st) {
  const PlaneEntry* e = plane_entry(s);
  FC_REQUIRE(e != nullptr, kInternal, "no whole-plane kernels for this plane extent");
  (dir < 0 ? e->fwd : e->inv)(a, st);
}

}  // namespace fcamd
