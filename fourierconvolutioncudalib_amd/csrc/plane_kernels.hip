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
// plane extents (nx, ny) with whole-plane kernels; each needs its
// half-spectrum plane and its row exchange in one workgroup's LDS
const PlaneEntry kPlanes[] = {FC_PLANE(64, 64), FC_PLANE(96, 96), FC_PLANE(128, 128),
                              FC_PLANE(128, 64), FC_PLANE(64, 128)};
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
