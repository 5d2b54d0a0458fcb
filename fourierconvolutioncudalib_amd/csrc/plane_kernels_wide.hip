// plane_kernels_wide.hip -- instantiations of the wide split plane kernels
// (plane_kernels.hpp): planes whose half-spectrum plane only fits one
// workgroup's LDS by halves (35-150 KB per part).
#include "plane_kernels.hpp"
#include "plane_registry.hpp"

namespace fcamd {

#define FC_PLANE_WIDE(NX, NY)                                                                      \
  {NX, NY, nullptr, nullptr, &plane_launch_fwd2w<NX, NY>, &plane_launch_inv2w<NX, NY>, true, \
   &plane_occupancy_fwd2w<NX, NY>, &plane_occupancy_inv2w<NX, NY>}
#define FC_PLANE_WIDE_INV(NX, NY)                                                               \
  {NX, NY, nullptr, nullptr, nullptr, &plane_launch_inv2w<NX, NY>, true, nullptr, \
   &plane_occupancy_inv2w<NX, NY>}
const PlaneEntry* wide_plane_entry(int nx, int ny) {
  // a host function's table: a namespace-scope one would also be emitted
  // into the device code, which has none of the host launchers
  static const PlaneEntry table[] = {
      // 240 x 240 keeps the separate forward passes: its wide forward
      // measured slower at 96 planes (57.8 -> 61.4 us per convolution),
      // where the other wide planes gained at 64 planes (256 x 240: 72.1 ->
      // 64.5 us forward split, 55.8 us both).
      FC_PLANE_WIDE(128, 160), FC_PLANE_WIDE(160, 128), FC_PLANE_WIDE(144, 144),
      FC_PLANE_WIDE(160, 160), FC_PLANE_WIDE(128, 192), FC_PLANE_WIDE(192, 128),
      FC_PLANE_WIDE(160, 192), FC_PLANE_WIDE(192, 160), FC_PLANE_WIDE(192, 192),
      FC_PLANE_WIDE(224, 224), FC_PLANE_WIDE_INV(240, 240), FC_PLANE_WIDE(256, 256),
      FC_PLANE_WIDE(128, 256), FC_PLANE_WIDE(256, 128), FC_PLANE_WIDE(192, 256),
      FC_PLANE_WIDE(256, 192), FC_PLANE_WIDE(224, 256), FC_PLANE_WIDE(256, 224),
      FC_PLANE_WIDE(240, 256), FC_PLANE_WIDE(256, 240),
  };
  for (const PlaneEntry& e : table)
    if (e.nx == nx && e.ny == ny) return &e;
  return nullptr;
}
#undef FC_PLANE_WIDE_INV
#undef FC_PLANE_WIDE

}  // namespace fcamd
