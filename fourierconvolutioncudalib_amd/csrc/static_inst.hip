// static_inst.hip -- instantiates the compile-time kernels for one group of
// lengths (FC_GROUP, set by the Makefile) and registers them on request.
// The groups split the instantiations over several objects so they compile
// in parallel.
#include "rl_kernels.hpp"
#include "static_kernels_impl.hpp"

#ifndef FC_GROUP
#error "FC_GROUP must be defined"
#endif

// Lengths with compile-time kernels: the image extents of the benchmark
// workloads and the uneven-block menu, plus common microscopy sizes.
#if FC_GROUP == 0
#define FC_LENGTHS(X) X(512) X(256)
#elif FC_GROUP == 1
#define FC_LENGTHS(X) X(600) X(500) X(240)
#elif FC_GROUP == 2
#define FC_LENGTHS(X) X(384) X(192) X(128) X(1024)
#elif FC_GROUP == 3
#define FC_LENGTHS(X) X(64) X(80) X(96) X(105) X(112)
#elif FC_GROUP == 4
#define FC_LENGTHS(X) X(120) X(144) X(160) X(210) X(224)
#elif FC_GROUP == 5
#define FC_LENGTHS(X) X(32) X(48) X(288) X(320) X(400) X(448) X(640) X(768)
#elif FC_GROUP == 6
#define FC_LENGTHS(X) X(100) X(150) X(200) X(250) X(286) X(300) X(360) X(480)
#elif FC_GROUP == 7
#define FC_LENGTHS(X) X(560) X(700) X(720) X(800) X(900) X(960) X(1000)
#elif FC_GROUP == 8
#define FC_LENGTHS(X) X(180) X(216) X(270) X(336) X(432) X(540) X(576) X(672) X(864) X(896)
#elif FC_GROUP == 9
#define FC_LENGTHS(X) X(1152) X(1200) X(1280) X(1536) X(2048)
#elif FC_GROUP == 10
#define FC_LENGTHS(X) X(2304) X(2560) X(3072) X(4096)
#else
#define FC_LENGTHS(X)
#endif

namespace fcamd {

#define FC_CAT2(a, b) a##b
#define FC_CAT(a, b) FC_CAT2(a, b)

// Registers this group's lengths.  Called by the registry on first lookup
// (static_kernels.hip), so a program that links the static archive pulls
// every group in through this reference; no load-time constructor is
// relied on.
void FC_CAT(static_register_group_, FC_GROUP)() {
#define FC_REG(L)                                \
  {                                              \
    StaticEntry e = make_static_entry<L>();      \
    e.rl.zfan = &rl_launch_zfan<L>;              \
    e.rl.zsum = &rl_launch_zsum<L>;              \
    e.rl.xratio = &rl_launch_xratio<L>;          \
    e.rl.xupdate = &rl_launch_xupdate<L>;        \
    static_register(L, e);                       \
  }
  FC_LENGTHS(FC_REG)
#undef FC_REG
}

}  // namespace fcamd
