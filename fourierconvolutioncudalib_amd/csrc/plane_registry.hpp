// plane_registry.hpp -- the table entry of one plane extent's plane kernels
// (plane_kernels.hip, plane_kernels_wide.hip).
#pragma once

#include "kernels.hpp"

namespace fcamd {

using PlaneFn = void (*)(const PlaneArgs&, hipStream_t);
using OccFn = int (*)();
struct PlaneEntry {
  int nx, ny;
  PlaneFn fwd, inv;          // one workgroup per plane (null for wide planes)
  PlaneFn fwd2, inv2;        // split kernels, where the half lengths have plans
  bool wide;                 // only the wide split kernels
  OccFn occ_fwd2, occ_inv2;  // wide: workgroups per CU of fwd2 / inv2
};

// entry of a wide plane extent (plane_kernels_wide.hip), or nullptr
const PlaneEntry* wide_plane_entry(int nx, int ny);

}  // namespace fcamd
