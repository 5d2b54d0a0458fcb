// batch.cpp -- many convolutions of uneven shapes in one plan.
//
// Every block (image, kernel, output; its own shape) goes through the same
// five passes as a single convolution.  Two schedules:
//
// Lanes (default).  The blocks are dealt to a few streams ("lanes"),
// largest estimated cost first, each to the lane with the least work so
// far; a lane runs its blocks' pass sequences one after another in its own
// workspace slot.  One block's spectrum (tens of MB) stays in the 256 MB
// Infinity Cache from pass to pass, and the lanes' launches overlap each
// other's ramp-up and drain, so neither HBM round trips between passes nor
// per-launch fixed costs are paid in full.  The workspace is one slot per
// lane.
//
// Length groups.  For each pass the blocks are grouped by the transform
// length of that pass's axis; each group is one launch of the compile-time
// kernel for that length over a device task table {argument record, tile},
// so a pass costs one launch per distinct length instead of one per block.
// Groups of one pass run on forked streams; the pass boundary joins them.
// Blocks run in waves: consecutive blocks whose workspaces add up to at
// most a byte budget run all passes before the next wave starts, sharing
// one workspace region.  Blocks whose length has no compile-time kernel
// fall back to the runtime-length kernels, one launch per block.
//
// Either sequence can be captured once and replayed: one hipGraph per lane
// (launched on the lane streams), or one graph for the length groups.
#include <algorithm>
#include <cstdlib>
#include <map>
#include <memory>
#include <vector>

#include "../../include/fourierconvolution_amd.h"
#include "conv_driver.hpp"
#include "static_registry.hpp"

namespace fcamd {

namespace {

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Default wave budget (bytes of block workspace per wave; 0 = one wave).
constexpr size_t kDefaultWaveBytes = 0;

// Default number of lanes (0 = the length-group schedule).  Four matches
// the hardware queues a process gets by default; more lanes share them.
constexpr int kDefaultLanes = 4;

// A block's estimated cost for lane balancing, in voxel units: its voxels
// plus a constant for the fixed cost of its five launches (~5 x 7 us, the
// time the passes take to move ~3.5 M voxels).
constexpr long kLaunchCostVoxels = 3500000;

struct Block {
  ConvShape s;
  float* out;
  const float* img;
  const float* ker;
  float2* S;
  float2* Kx;
  float2* Kxy;
};

template <class A>
struct Group {
  int L = 0;
  int path = 0;  // z pass: 1 = the two-stage tap path (all kz <= sparse_kz)
  StaticEntry* e = nullptr;
  std::vector<A> args;
  std::vector<int2> tasks;
  A* d_args = nullptr;
  int2* d_tasks = nullptr;
};

template <class A>
void upload(Group<A>& g) {
  FC_HIP_CHECK(hipMalloc(&g.d_args, sizeof(A) * std::max<size_t>(1, g.args.size())));
  FC_HIP_CHECK(hipMalloc(&g.d_tasks, sizeof(int2) * std::max<size_t>(1, g.tasks.size())));
  FC_HIP_CHECK(hipMemcpy(g.d_args, g.args.data(), sizeof(A) * g.args.size(), hipMemcpyHostToDevice));
  FC_HIP_CHECK(hipMemcpy(g.d_tasks, g.tasks.data(), sizeof(int2) * g.tasks.size(),
                         hipMemcpyHostToDevice));
}

template <class A>
void release(Group<A>& g) {
  if (g.d_args) (void)hipFree(g.d_args);
  if (g.d_tasks) (void)hipFree(g.d_tasks);
  g.d_args = nullptr;
  g.d_tasks = nullptr;
}

void add_tasks(std::vector<int2>& tasks, int arg, long count) {
  FC_REQUIRE(count < (1L << 31), kUnsupportedSize, "batch block too large for the task table");
  for (long i = 0; i < count; ++i) tasks.push_back(make_int2(arg, (int)i));
}

}  // namespace

// Wave partition of n blocks with workspace sizes ws[b]: consecutive blocks
// while the running sum stays within `budget` (0 = one wave); returns the
// first block of every wave and the largest wave's workspace.
std::vector<int> wave_starts(const std::vector<size_t>& ws, size_t budget, size_t* largest) {
  std::vector<int> starts;
  size_t cur = 0, big = 0;
  for (int b = 0; b < (int)ws.size(); ++b) {
    if (b == 0 || (budget > 0 && cur + ws[b] > budget)) {
      starts.push_back(b);
      cur = 0;
    }
    cur += ws[b];
    big = std::max(big, cur);
  }
  if (largest) *largest = big;
  return starts;
}

// Wave byte budget: FCAMD_BATCH_WAVE_MB when set (0 = one wave), else the
// default below.
size_t default_wave_budget() {
  if (const char* v = std::getenv("FCAMD_BATCH_WAVE_MB")) return (size_t)std::atol(v) << 20;
  return kDefaultWaveBytes;
}

// Blocks the drop-in call would send to the long-axis path (axes past one
// tile, large prime factors) run there in the lane schedule; the length
// groups take one-tile blocks only.
size_t block_workspace(const ConvShape& s) {
  return align256(needs_long_path(s) ? long_workspace_bytes(s) : workspace_layout(s).total());
}

void validate_block(const ConvShape& s, int nlanes) {
  if (!needs_long_path(s)) {
    validate_shape(s);
    return;
  }
  validate_long_shape(s);
  FC_REQUIRE(nlanes > 0, kUnsupportedSize,
             "a block needs the long-axis path, which the length-group schedule "
             "(FCAMD_BATCH_LANES=0) does not run");
}

void enqueue_block(const ConvShape& s, float* out, const float* img, const float* ker, Op op,
                   void* ws, hipStream_t st) {
  if (needs_long_path(s))
    enqueue_convolution_long(s, out, img, ker, op, ws, st);
  else
    enqueue_convolution(s, out, img, ker, op, ws, st);
}

// Lane count: FCAMD_BATCH_LANES when set (0 = length groups), else the
// default above.
int default_lanes() {
  if (const char* v = std::getenv("FCAMD_BATCH_LANES")) return std::max(0, std::atoi(v));
  return kDefaultLanes;
}

// Lane assignment (longest processing time first): blocks by decreasing
// estimated cost, each to the lane with the least accumulated cost; ties
// go to the lowest lane.  Returns the blocks of every non-empty lane in
// execution order and each lane's workspace slot (its largest block).
std::vector<std::vector<int>> lane_plan(const std::vector<ConvShape>& shapes, int nlanes,
                                        std::vector<size_t>* slot_bytes) {
  const int n = (int)shapes.size();
  nlanes = std::max(1, std::min(nlanes, n));
  std::vector<int> order(n);
  std::vector<long> cost(n);
  for (int b = 0; b < n; ++b) {
    order[b] = b;
    cost[b] = shapes[b].voxels() + kLaunchCostVoxels;
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
  std::vector<std::vector<int>> lanes(nlanes);
  std::vector<long> load(nlanes, 0);
  std::vector<size_t> slot(nlanes, 0);
  for (int b : order) {
    const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    lanes[l].push_back(b);
    load[l] += cost[b];
    slot[l] = std::max(slot[l], block_workspace(shapes[b]));
  }
  if (slot_bytes) *slot_bytes = slot;
  return lanes;
}

ConvShape block_shape(const int* imDims, const int* kernelDims, int b) {
  ConvShape s{};
  s.nx = imDims[3 * b];
  s.ny = imDims[3 * b + 1];
  s.nz = imDims[3 * b + 2];
  s.kx = kernelDims[3 * b];
  s.ky = kernelDims[3 * b + 1];
  s.kz = kernelDims[3 * b + 2];
  s.cx = s.kx / 2;
  s.cy = s.ky / 2;
  s.cz = s.kz / 2;
  return s;
}

class BatchPlan {
 public:
  BatchPlan(int n, const int* imDims, const int* kernelDims, float* const* outs,
            const float* const* imgs, const float* const* kers, int op, void* ws, size_t ws_bytes,
            size_t wave_budget, int nlanes);
  ~BatchPlan();
  void enqueue(hipStream_t st);
  void graph_launch(hipStream_t st);
  size_t workspace() const { return ws_need_; }

 private:
  struct Wave {
    std::vector<Group<XPassArgs>> xf, xi;
    std::vector<Group<ColArgs>> yf, yi;
    std::vector<Group<ZArgs>> z;
    std::vector<int> generic_x, generic_y, generic_z;  // block indices
  };
  template <class F>
  void fork(hipStream_t st, size_t ngroups, F launch_group);
  void build_wave(Wave& w, int first, int last);
  void enqueue_wave(const Wave& w, hipStream_t st);
  void enqueue_lanes(hipStream_t st);

  int dev_ = 0;
  bool conj_ = false;
  std::vector<Block> blocks_;
  size_t ws_need_ = 0;
  void* ws_ = nullptr;
  bool own_ws_ = false;
  std::vector<Wave> waves_;
  // lane schedule: blocks of every lane, and each lane's workspace slot
  std::vector<std::vector<int>> lanes_;
  std::vector<size_t> lane_off_;
  std::vector<hipStream_t> side_;
  std::vector<hipEvent_t> ev_;
  hipGraphExec_t exec_ = nullptr;
  hipGraph_t graph_ = nullptr;
  hipStream_t cap_ = nullptr;
  // lane schedule: one captured graph per lane, launched on the lane's stream
  std::vector<hipGraph_t> lane_graphs_;
  std::vector<hipGraphExec_t> lane_execs_;
  std::vector<hipStream_t> lane_caps_;
};

BatchPlan::BatchPlan(int n, const int* imDims, const int* kernelDims, float* const* outs,
                     const float* const* imgs, const float* const* kers, int op, void* ws,
                     size_t ws_bytes, size_t wave_budget, int nlanes) {
  FC_REQUIRE(n >= 1, kInvalidArgument, "empty batch");
  FC_REQUIRE(imDims && kernelDims && outs && imgs && kers, kInvalidArgument, "null array");
  FC_REQUIRE(op == FCAMD_CONVOLVE || op == FCAMD_CORRELATE, kInvalidArgument, "unknown op");
  conj_ = op == FCAMD_CORRELATE;
  FC_HIP_CHECK(hipGetDevice(&dev_));
  std::vector<size_t> sizes;
  for (int b = 0; b < n; ++b) {
    Block bl{};
    bl.s = block_shape(imDims, kernelDims, b);
    validate_block(bl.s, nlanes);
    FC_REQUIRE(outs[b] && imgs[b] && kers[b], kInvalidArgument, "null block pointer");
    bl.out = outs[b];
    bl.img = imgs[b];
    bl.ker = kers[b];
    sizes.push_back(block_workspace(bl.s));
    blocks_.push_back(bl);
  }
  std::vector<int> starts;
  if (nlanes > 0) {
    std::vector<ConvShape> shapes;
    for (const Block& bl : blocks_) shapes.push_back(bl.s);
    std::vector<size_t> slot;
    lanes_ = lane_plan(shapes, nlanes, &slot);
    ws_need_ = 0;
    for (size_t l = 0; l < lanes_.size(); ++l) {
      lane_off_.push_back(ws_need_);
      ws_need_ += slot[l];
    }
  } else {
    starts = wave_starts(sizes, wave_budget, &ws_need_);
  }
  if (ws) {
    FC_REQUIRE(ws_bytes >= ws_need_, kInvalidArgument, "batch workspace too small");
    ws_ = ws;
  } else {
    FC_HIP_CHECK(hipMalloc(&ws_, ws_need_));
    own_ws_ = true;
  }
  driver_init();
  // create every twiddle table now: nothing may allocate or copy while a
  // graph is being captured
  for (const Block& bl : blocks_) {
    if (needs_long_path(bl.s)) {
      long_prepare_tables(bl.s, nullptr);
      continue;
    }
    int lens[5];
    for (int i = 0, n = table_lengths(bl.s, lens); i < n; ++i) (void)fft_desc(lens[i]);
  }
  if (nlanes > 0) {
    for (size_t i = 0; i + 1 < lanes_.size(); ++i) {
      hipStream_t st = nullptr;
      FC_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      side_.push_back(st);
    }
    for (size_t i = 0; i < lanes_.size(); ++i) {
      hipEvent_t e = nullptr;
      FC_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      ev_.push_back(e);
    }
    return;
  }
  // every wave lays its blocks out from the start of the workspace
  for (size_t w = 0; w < starts.size(); ++w) {
    const int last = w + 1 < starts.size() ? starts[w + 1] : n;
    size_t off = 0;
    for (int b = starts[w]; b < last; ++b) {
      const WorkspaceLayout wl = workspace_layout(blocks_[b].s);
      char* base = static_cast<char*>(ws_) + off;
      blocks_[b].S = reinterpret_cast<float2*>(base);
      blocks_[b].Kx = reinterpret_cast<float2*>(base + wl.spectrum);
      blocks_[b].Kxy = reinterpret_cast<float2*>(base + wl.spectrum + wl.krows);
      off += sizes[b];
    }
  }
  waves_.resize(starts.size());
  size_t nside = 1;
  for (size_t w = 0; w < starts.size(); ++w) {
    build_wave(waves_[w], starts[w], w + 1 < starts.size() ? starts[w + 1] : n);
    nside = std::max({nside, waves_[w].xf.size(), waves_[w].yf.size(), waves_[w].z.size()});
  }
  for (size_t i = 0; i < nside; ++i) {
    hipStream_t st = nullptr;
    FC_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    side_.push_back(st);
  }
  for (size_t i = 0; i < nside + 1; ++i) {
    hipEvent_t e = nullptr;
    FC_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    ev_.push_back(e);
  }
}

void BatchPlan::build_wave(Wave& w, int first, int last) {
  std::map<int, std::vector<int>> by_x, by_y, by_z;
  for (int b = first; b < last; ++b) {
    by_x[blocks_[b].s.nx].push_back(b);
    by_y[blocks_[b].s.ny].push_back(b);
    by_z[blocks_[b].s.nz].push_back(b);
  }
  // x forward / inverse
  for (auto& kv : by_x) {
    StaticEntry* e = static_lookup(kv.first);
    if (!e || !e->batch.xfwd) {
      w.generic_x.insert(w.generic_x.end(), kv.second.begin(), kv.second.end());
      continue;
    }
    const FftDesc fx = fft_desc(kv.first);
    const int P = e->batch.pairs;
    Group<XPassArgs> gf, gi;
    gf.L = gi.L = kv.first;
    gf.e = gi.e = e;
    for (int b : kv.second) {
      const Block& bl = blocks_[b];
      const long nxp = padded_bins(bl.s.nx);
      XPassArgs m{};
      m.fft = fx;
      m.nrows = (long)bl.s.ny * bl.s.nz;
      m.nbins = half_bins(bl.s.nx);
      m.src = bl.img;
      m.dst = bl.S;
      m.src_row = bl.s.nx;
      m.dst_row = nxp;
      m.klen = bl.s.nx;
      m.kcenter = 0;
      XPassArgs k = m;
      k.nrows = (long)bl.s.ky * bl.s.kz;
      k.src = bl.ker;
      k.dst = bl.Kx;
      k.src_row = bl.s.kx;
      k.klen = bl.s.kx;
      k.kcenter = bl.s.cx;
      k.kshift = 1;
      // kernel rows first (they gate the kernel planes of the next pass)
      gf.args.push_back(k);
      add_tasks(gf.tasks, (int)gf.args.size() - 1, ((k.nrows + 1) / 2 + P - 1) / P);
      gf.args.push_back(m);
      add_tasks(gf.tasks, (int)gf.args.size() - 1, ((m.nrows + 1) / 2 + P - 1) / P);
      XPassArgs iv{};
      iv.fft = fx;
      iv.nrows = m.nrows;
      iv.nbins = m.nbins;
      iv.csrc = bl.S;
      iv.rdst = bl.out;
      iv.src_row = nxp;
      iv.dst_row = bl.s.nx;
      gi.args.push_back(iv);
      add_tasks(gi.tasks, (int)gi.args.size() - 1, ((iv.nrows + 1) / 2 + P - 1) / P);
    }
    upload(gf);
    upload(gi);
    w.xf.push_back(gf);
    w.xi.push_back(gi);
  }
  // y forward (+ kernel planes) / inverse
  for (auto& kv : by_y) {
    StaticEntry* e = static_lookup(kv.first);
    if (!e || !e->batch.col) {
      w.generic_y.insert(w.generic_y.end(), kv.second.begin(), kv.second.end());
      continue;
    }
    const FftDesc fy = fft_desc(kv.first);
    const int TX = e->batch.tx;
    Group<ColArgs> gf, gi;
    gf.L = gi.L = kv.first;
    gf.e = gi.e = e;
    for (int b : kv.second) {
      const Block& bl = blocks_[b];
      const long nxp = padded_bins(bl.s.nx);
      const int nb = half_bins(bl.s.nx);
      ColArgs m{};
      m.fft = fy;
      m.nbins = nb;
      m.src = bl.S;
      m.dst = bl.S;
      m.src_outer = m.dst_outer = (long)bl.s.ny * nxp;
      m.src_elem = m.dst_elem = nxp;
      m.nvalid = bl.s.ny;
      m.center = 0;
      m.nouter = bl.s.nz;
      ColArgs k = m;
      k.src = bl.Kx;
      k.dst = bl.Kxy;
      k.src_outer = (long)bl.s.ky * nxp;
      k.nvalid = bl.s.ky;
      k.center = bl.s.cy;
      k.nouter = bl.s.kz;
      gf.args.push_back(k);
      add_tasks(gf.tasks, (int)gf.args.size() - 1, column_tiles(nb, TX, bl.s.kz));
      gf.args.push_back(m);
      add_tasks(gf.tasks, (int)gf.args.size() - 1, column_tiles(nb, TX, bl.s.nz));
      gi.args.push_back(m);
      add_tasks(gi.tasks, (int)gi.args.size() - 1, column_tiles(nb, TX, bl.s.nz));
    }
    upload(gf);
    upload(gi);
    w.yf.push_back(gf);
    w.yi.push_back(gi);
  }
  // fused z
  for (auto& kv : by_z) {
    StaticEntry* e = static_lookup(kv.first);
    if (!e || !e->batch.zfused) {
      w.generic_z.insert(w.generic_z.end(), kv.second.begin(), kv.second.end());
      continue;
    }
    const FftDesc fz = fft_desc(kv.first);
    const int TX = e->batch.tx;
    // blocks whose kz fits the tap path form their own launch
    Group<ZArgs> gs[2];
    for (int p = 0; p < 2; ++p) {
      gs[p].L = kv.first;
      gs[p].e = e;
      gs[p].path = p;
    }
    for (int b : kv.second) {
      Group<ZArgs>& g = gs[e->batch.zfused_sparse && blocks_[b].s.kz <= e->batch.sparse_kz ? 1 : 0];
      const Block& bl = blocks_[b];
      const long nxp = padded_bins(bl.s.nx);
      ZArgs a{};
      a.fft = fz;
      a.nbins = half_bins(bl.s.nx);
      a.s = bl.S;
      a.s_row = nxp;
      a.s_plane = (long)bl.s.ny * nxp;
      a.k = bl.Kxy;
      a.k_row = nxp;
      a.k_plane = (long)bl.s.ny * nxp;
      a.nvalid = bl.s.kz;
      a.center = bl.s.cz;
      a.conj = conj_ ? 1 : 0;
      a.scale = (float)(1.0 / (double)bl.s.voxels());
      a.ny = bl.s.ny;
      g.args.push_back(a);
      add_tasks(g.tasks, (int)g.args.size() - 1, column_tiles(a.nbins, TX, bl.s.ny));
    }
    for (Group<ZArgs>& g : gs) {
      if (g.args.empty()) continue;
      upload(g);
      w.z.push_back(g);
    }
  }
}

BatchPlan::~BatchPlan() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  for (auto x : lane_execs_) (void)hipGraphExecDestroy(x);
  for (auto g : lane_graphs_) (void)hipGraphDestroy(g);
  for (auto c : lane_caps_) (void)hipStreamDestroy(c);
  for (Wave& w : waves_) {
    for (auto& g : w.xf) release(g);
    for (auto& g : w.xi) release(g);
    for (auto& g : w.yf) release(g);
    for (auto& g : w.yi) release(g);
    for (auto& g : w.z) release(g);
  }
  for (auto s : side_) (void)hipStreamDestroy(s);
  if (cap_) (void)hipStreamDestroy(cap_);
  for (auto e : ev_) (void)hipEventDestroy(e);
  if (own_ws_ && ws_) (void)hipFree(ws_);
}

// Runs launch_group(i, stream) for every group i, group 0 on `st` and the
// others on side streams forked from and joined back into `st`.
template <class F>
void BatchPlan::fork(hipStream_t st, size_t ngroups, F launch_group) {
  if (ngroups == 0) return;
  if (ngroups > 1) {
    FC_HIP_CHECK(hipEventRecord(ev_[0], st));
    for (size_t i = 1; i < ngroups; ++i) FC_HIP_CHECK(hipStreamWaitEvent(side_[i - 1], ev_[0], 0));
  }
  for (size_t i = 0; i < ngroups; ++i) launch_group(i, i == 0 ? st : side_[i - 1]);
  for (size_t i = 1; i < ngroups; ++i) {
    FC_HIP_CHECK(hipEventRecord(ev_[i], side_[i - 1]));
    FC_HIP_CHECK(hipStreamWaitEvent(st, ev_[i], 0));
  }
}

void BatchPlan::enqueue_wave(const Wave& w, hipStream_t st) {
  // x forward (image and kernel rows)
  fork(st, w.xf.size(), [&](size_t i, hipStream_t s) {
    auto& g = w.xf[i];
    g.e->batch.xfwd(g.d_args, g.d_tasks, (int)g.tasks.size(), s);
  });
  for (int b : w.generic_x) {
    const Block& bl = blocks_[b];
    pass_x_forward(bl.s, bl.img, bl.S, bl.ker, bl.Kx, st);
  }
  // y forward (image and kernel planes)
  fork(st, w.yf.size(), [&](size_t i, hipStream_t s) {
    auto& g = w.yf[i];
    g.e->batch.col(g.d_args, g.d_tasks, (int)g.tasks.size(), -1, s);
  });
  for (int b : w.generic_y) {
    const Block& bl = blocks_[b];
    pass_y_forward(bl.s, bl.S, bl.Kx, bl.Kxy, st);
  }
  fork(st, w.z.size(), [&](size_t i, hipStream_t s) {
    auto& g = w.z[i];
    (g.path == 1 ? g.e->batch.zfused_sparse : g.e->batch.zfused)(g.d_args, g.d_tasks,
                                                                 (int)g.tasks.size(), s);
  });
  for (int b : w.generic_z) {
    const Block& bl = blocks_[b];
    z_pass(bl.s, bl.S, bl.Kxy, false, conj_, st);
  }
  fork(st, w.yi.size(), [&](size_t i, hipStream_t s) {
    auto& g = w.yi[i];
    g.e->batch.col(g.d_args, g.d_tasks, (int)g.tasks.size(), +1, s);
  });
  for (int b : w.generic_y) pass_y_inverse(blocks_[b].s, blocks_[b].S, st);
  fork(st, w.xi.size(), [&](size_t i, hipStream_t s) {
    auto& g = w.xi[i];
    g.e->batch.xinv(g.d_args, g.d_tasks, (int)g.tasks.size(), s);
  });
  for (int b : w.generic_x) pass_x_inverse(blocks_[b].s, blocks_[b].S, blocks_[b].out, st);
}

// Lane 0 runs on `st`, the others on side streams forked from and joined
// back into `st`.  Blocks are enqueued round-robin over the lanes, so in
// eager mode every lane has work queued from the start.
void BatchPlan::enqueue_lanes(hipStream_t st) {
  const size_t nl = lanes_.size();
  if (nl > 1) {
    FC_HIP_CHECK(hipEventRecord(ev_[0], st));
    for (size_t i = 1; i < nl; ++i) FC_HIP_CHECK(hipStreamWaitEvent(side_[i - 1], ev_[0], 0));
  }
  const Op op = conj_ ? Op::kCorrelate : Op::kConvolve;
  size_t depth = 0;
  for (const auto& lane : lanes_) depth = std::max(depth, lane.size());
  for (size_t j = 0; j < depth; ++j) {
    for (size_t i = 0; i < nl; ++i) {
      if (j >= lanes_[i].size()) continue;
      const Block& bl = blocks_[lanes_[i][j]];
      enqueue_block(bl.s, bl.out, bl.img, bl.ker, op, static_cast<char*>(ws_) + lane_off_[i],
                    i == 0 ? st : side_[i - 1]);
    }
  }
  for (size_t i = 1; i < nl; ++i) {
    FC_HIP_CHECK(hipEventRecord(ev_[i], side_[i - 1]));
    FC_HIP_CHECK(hipStreamWaitEvent(st, ev_[i], 0));
  }
}

void BatchPlan::enqueue(hipStream_t st) {
  if (!lanes_.empty()) {
    enqueue_lanes(st);
  } else {
    // waves share the workspace: stream order on `st` (every fork joins
    // back into it) keeps wave w+1 from starting before wave w has finished
    for (const Wave& w : waves_) enqueue_wave(w, st);
  }
  FC_HIP_CHECK(hipGetLastError());
}

// Lane schedule: one graph per lane, each a linear chain, launched on the
// lane's stream (1-2 % faster than one graph holding all lanes, whose
// replay the runtime schedules itself).  Length groups: one graph.
void BatchPlan::graph_launch(hipStream_t st) {
  if (!lanes_.empty()) {
    const size_t nl = lanes_.size();
    if (lane_execs_.size() != nl) {
      // each lane's pass sequence captured on a plan-owned stream; the
      // graphs are kept only once every lane has been captured and
      // instantiated, so a failure leaves no partial set behind
      const Op op = conj_ ? Op::kCorrelate : Op::kConvolve;
      while (lane_caps_.size() < nl) {
        hipStream_t cs = nullptr;
        FC_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        lane_caps_.push_back(cs);
      }
      std::vector<hipGraph_t> graphs;
      std::vector<hipGraphExec_t> execs;
      auto drop = [&] {
        for (auto x : execs) (void)hipGraphExecDestroy(x);
        for (auto g : graphs) (void)hipGraphDestroy(g);
      };
      try {
        for (size_t i = 0; i < nl; ++i) {
          hipStream_t cs = lane_caps_[i];
          FC_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
          try {
            for (int b : lanes_[i]) {
              const Block& bl = blocks_[b];
              enqueue_block(bl.s, bl.out, bl.img, bl.ker, op,
                            static_cast<char*>(ws_) + lane_off_[i], cs);
            }
          } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(cs, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
          }
          hipGraph_t g = nullptr;
          FC_HIP_CHECK(hipStreamEndCapture(cs, &g));
          graphs.push_back(g);
          hipGraphExec_t x = nullptr;
          FC_HIP_CHECK(hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
          execs.push_back(x);
        }
      } catch (...) {
        drop();
        throw;
      }
      lane_graphs_ = std::move(graphs);
      lane_execs_ = std::move(execs);
    }
    // lane 0 on `st`, the others on side streams forked from and joined
    // back into `st`
    if (nl > 1) {
      FC_HIP_CHECK(hipEventRecord(ev_[0], st));
      for (size_t i = 1; i < nl; ++i) FC_HIP_CHECK(hipStreamWaitEvent(side_[i - 1], ev_[0], 0));
    }
    for (size_t i = 0; i < nl; ++i)
      FC_HIP_CHECK(hipGraphLaunch(lane_execs_[i], i == 0 ? st : side_[i - 1]));
    for (size_t i = 1; i < nl; ++i) {
      FC_HIP_CHECK(hipEventRecord(ev_[i], side_[i - 1]));
      FC_HIP_CHECK(hipStreamWaitEvent(st, ev_[i], 0));
    }
    return;
  }
  if (!exec_) {
    // capture on a plan-owned stream (the caller's may be the null stream,
    // which cannot be captured); the graph is then launched on `st`
    if (!cap_) FC_HIP_CHECK(hipStreamCreateWithFlags(&cap_, hipStreamNonBlocking));
    FC_HIP_CHECK(hipStreamBeginCapture(cap_, hipStreamCaptureModeThreadLocal));
    try {
      enqueue(cap_);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(cap_, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    hipGraph_t g = nullptr;
    FC_HIP_CHECK(hipStreamEndCapture(cap_, &g));
    hipGraphExec_t x = nullptr;
    const hipError_t rc = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (rc != hipSuccess) {
      (void)hipGraphDestroy(g);
      FC_HIP_CHECK(rc);
    }
    graph_ = g;
    exec_ = x;
  }
  FC_HIP_CHECK(hipGraphLaunch(exec_, st));
}

}  // namespace fcamd

// ===========================================================================
// C ABI
// ===========================================================================
using namespace fcamd;

namespace {
template <class F>
int batch_guarded(F&& f) {
  clear_error();
  try {
    f();
    return kOk;
  } catch (const Error& e) {
    set_error(e.code, e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error(kOutOfMemory, "host allocation failed");
    return kOutOfMemory;
  } catch (...) {
    set_error(kInternal, "unexpected internal error");
    return kInternal;
  }
}
}  // namespace

extern "C" {

long long fcamd_batch_workspace_bytes(int nblocks, const int* imDims, const int* kernelDims) {
  long long r = -1;
  (void)batch_guarded([&] {
    FC_REQUIRE(nblocks >= 1 && imDims && kernelDims, kInvalidArgument, "bad batch description");
    std::vector<size_t> sizes;
    std::vector<ConvShape> shapes;
    for (int b = 0; b < nblocks; ++b) {
      shapes.push_back(block_shape(imDims, kernelDims, b));
    }
    const int nlanes = default_lanes();
    for (const ConvShape& s : shapes) {
      validate_block(s, nlanes);
      sizes.push_back(block_workspace(s));
    }
    size_t need = 0;
    if (nlanes > 0) {
      std::vector<size_t> slot;
      (void)lane_plan(shapes, nlanes, &slot);
      for (size_t b : slot) need += b;
    } else {
      (void)wave_starts(sizes, default_wave_budget(), &need);
    }
    r = (long long)need;
  });
  return r;
}

void* fcamd_batch_create(int nblocks, const int* imDims, const int* kernelDims, float* const* d_outs,
                         const float* const* d_images, const float* const* d_kernels, int op,
                         void* d_workspace, long long workspace_bytes) {
  BatchPlan* p = nullptr;
  (void)batch_guarded([&] {
    p = new BatchPlan(nblocks, imDims, kernelDims, d_outs, d_images, d_kernels, op, d_workspace,
                      (size_t)(workspace_bytes < 0 ? 0 : workspace_bytes), default_wave_budget(),
                      default_lanes());
  });
  return p;
}

int fcamd_batch_execute(void* plan, void* stream) {
  return batch_guarded([&] {
    FC_REQUIRE(plan != nullptr, kInvalidArgument, "null batch plan");
    static_cast<BatchPlan*>(plan)->enqueue(static_cast<hipStream_t>(stream));
  });
}

int fcamd_batch_graph_launch(void* plan, void* stream) {
  return batch_guarded([&] {
    FC_REQUIRE(plan != nullptr, kInvalidArgument, "null batch plan");
    static_cast<BatchPlan*>(plan)->graph_launch(static_cast<hipStream_t>(stream));
  });
}

void fcamd_batch_destroy(void* plan) { delete static_cast<BatchPlan*>(plan); }

}  // extern "C"
