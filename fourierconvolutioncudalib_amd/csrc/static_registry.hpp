// static_registry.hpp -- lengths that have compile-time kernels, their
// variants (tile shape, pipelining, kernel-spectrum method) and launchers.
// Filled at load time by the instantiation units; the autotuner picks the
// variant per (pass, length).
#pragma once

#include <vector>

#include "kernels.hpp"

namespace fcamd {

// kSCol is the forward y pass, kSColInv the inverse one; both choose among
// the same column variants
enum StaticPass : int { kSxFwd = 0, kSxInv = 1, kSCol = 2, kSZ = 3, kSColInv = 4, kSNumPasses = 5 };

// Every launcher returns false when its variant does not apply to the
// arguments (the dispatcher then uses variant 0).
using XFwdFn = bool (*)(const XPass2&, hipStream_t);
using XInvFn = bool (*)(const XPassArgs&, hipStream_t);
using ColFn = bool (*)(const Col2&, int dir, hipStream_t);
using ZFn = bool (*)(const ZArgs&, long ny, hipStream_t);

template <class F>
struct Variant {
  const char* name;
  F fn;
  int tile = 0;  // column variants: bins per workgroup of a plain tile (0: other forms)
};

// batched forms: one launch over a task table {argument record, tile}
struct StaticBatch {
  int pairs = 1;  // row pairs per workgroup of the x passes
  int tx = 1;     // bins per tile of the column / z passes
  void (*xfwd)(const XPassArgs*, const int2*, int, hipStream_t) = nullptr;
  void (*xinv)(const XPassArgs*, const int2*, int, hipStream_t) = nullptr;
  void (*col)(const ColArgs*, const int2*, int, int dir, hipStream_t) = nullptr;
  void (*zfused)(const ZArgs*, const int2*, int, hipStream_t) = nullptr;
  // the two-stage tap path; every task's kz must be <= sparse_kz
  void (*zfused_sparse)(const ZArgs*, const int2*, int, hipStream_t) = nullptr;
  int sparse_kz = 0;
};

// fused Richardson-Lucy kernels (rl_kernels.hpp)
struct StaticRL {
  void (*zfan)(const RLZArgs&, long ny, hipStream_t) = nullptr;
  void (*zsum)(const RLZArgs&, long ny, hipStream_t) = nullptr;
  void (*xratio)(const RLXArgs&, int nviews, hipStream_t) = nullptr;
  void (*xupdate)(const RLXArgs&, hipStream_t) = nullptr;
};

struct StaticEntry {
  std::vector<Variant<XFwdFn>> xfwd;
  std::vector<Variant<XInvFn>> xinv;
  std::vector<Variant<ColFn>> col;
  std::vector<Variant<ZFn>> zfused;
  XFwdFn xfwd_vpad = nullptr;  // x forward over a virtually padded image
  XInvFn xinv_crop = nullptr;  // x inverse writing the cropped interior
  int selected[kSNumPasses] = {0, 0, 0, 0, 0};
  StaticBatch batch;
  StaticRL rl;
  int count(int pass) const {
    switch (pass) {
      case kSxFwd: return (int)xfwd.size();
      case kSxInv: return (int)xinv.size();
      case kSCol:
      case kSColInv: return (int)col.size();
      case kSZ: return (int)zfused.size();
      default: return 0;
    }
  }
  const char* name(int pass, int i) const {
    if (i < 0 || i >= count(pass)) return nullptr;
    switch (pass) {
      case kSxFwd: return xfwd[i].name;
      case kSxInv: return xinv[i].name;
      case kSCol:
      case kSColInv: return col[i].name;
      default: return zfused[i].name;
    }
  }
};

void static_register(int L, const StaticEntry& e);
// Entry for length L (nullptr when none or when compile-time kernels are
// switched off).  The pointer stays valid for the life of the process.
StaticEntry* static_lookup(int L);

// Persistent-grid size for a kernel: resident workgroups per CU x CUs.
int persistent_blocks(const void* kernel, int threads);

}  // namespace fcamd
