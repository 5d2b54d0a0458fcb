// static_registry.hpp -- lengths that have compile-time kernels, and their
// launchers.  Filled at load time by the instantiation units.
#pragma once

#include "kernels.hpp"

namespace fcamd {

struct StaticEntry {
  void (*xfwd)(const XPassArgs&, long npairs, hipStream_t) = nullptr;
  void (*xinv)(const XPassArgs&, long npairs, hipStream_t) = nullptr;
  void (*col)(const ColArgs&, int dir, long nouter, hipStream_t) = nullptr;
  void (*zfused)(const ZArgs&, long ny, hipStream_t) = nullptr;
};

void static_register(int L, const StaticEntry& e);
const StaticEntry* static_lookup(int L);

}  // namespace fcamd
