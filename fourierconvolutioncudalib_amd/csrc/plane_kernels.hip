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
// plane extents (nx, ny) with whole-plane kernels: every pair of the
// compile-time lengths 64-128 with nx a multiple of 4 (35-139 KB of LDS:
// the half-spectrum plane and the row exchange of one workgroup)
const PlaneEntry kPlanes[] = {FC_PLANE_ROW(64),  FC_PLANE_ROW(80),  FC_PLANE_ROW(96),
                              FC_PLANE_ROW(112), FC_PLANE_ROW(120), FC_PLANE_ROW(128)};
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
