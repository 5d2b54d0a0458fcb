// slabs.cpp -- host-pointer convolution streamed through the device in
// z-slabs (overlap-save).
//
// The output planes are cut into slabs of m planes.  Slab [z0, z0 + m) is
// the circular convolution of a virtually padded problem whose z axis reads
// the full volume's planes (z0 - lo + p) mod nz for p < P = fast length
// >= m + kz - 1 (lo = kz - 1 - kz/2), so its interior equals the circular
// convolution of the whole volume there; x and y are not padded.  The host
// volume is uploaded in plane chunks (tail planes needed by the first slab
// first), each slab's convolution starts as soon as its planes are on the
// device, and finished slabs are downloaded while later planes are still
// being uploaded: the two PCIe directions, the host-side staging copies and
// the device work overlap.  All transfers go through pinned buffers.
#include "slabs.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/fourierconvolution_amd.h"

namespace fcamd {
namespace {

constexpr size_t kSlabChunk = size_t(8) << 20;  // target bytes per transfer chunk
constexpr int kUpBufs = 4, kDownBufs = 4;

struct SlabStreams {
  hipStream_t up = nullptr, down = nullptr;
  std::vector<void*> ubuf, dbuf;  // pinned
  size_t buf_bytes = 0;
};

std::mutex g_ss_mu;
std::vector<std::pair<int, SlabStreams*>> g_ss;  // per device, kept for the process

SlabStreams& slab_streams(int dev, size_t buf_bytes) {
  std::lock_guard<std::mutex> lk(g_ss_mu);
  SlabStreams* s = nullptr;
  for (auto& e : g_ss)
    if (e.first == dev) s = e.second;
  if (!s) {
    s = new SlabStreams();
    FC_HIP_CHECK(hipStreamCreateWithFlags(&s->up, hipStreamNonBlocking));
    FC_HIP_CHECK(hipStreamCreateWithFlags(&s->down, hipStreamNonBlocking));
    g_ss.push_back({dev, s});
  }
  if (s->buf_bytes < buf_bytes) {
    for (void* p : s->ubuf) (void)hipHostFree(p);
    for (void* p : s->dbuf) (void)hipHostFree(p);
    s->ubuf.assign(kUpBufs, nullptr);
    s->dbuf.assign(kDownBufs, nullptr);
    s->buf_bytes = 0;
    for (auto* v : {&s->ubuf, &s->dbuf})
      for (void*& p : *v) FC_HIP_CHECK(hipHostMalloc(&p, buf_bytes, hipHostMallocDefault));
    s->buf_bytes = buf_bytes;
  }
  return *s;
}

struct EventPool {
  std::vector<hipEvent_t> ev;
  hipEvent_t get(size_t i) {
    while (ev.size() <= i) {
      hipEvent_t e = nullptr;
      FC_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      ev.push_back(e);
    }
    return ev[i];
  }
  ~EventPool() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  }
};

bool done(hipEvent_t e) {
  const hipError_t r = hipEventQuery(e);
  if (r == hipSuccess) return true;
  if (r == hipErrorNotReady) return false;
  FC_HIP_CHECK(r);
  return false;
}

}  // namespace

SlabPlan slab_plan(const ConvShape& s) {
  SlabPlan p{};
  const size_t plane = sizeof(float) * (size_t)s.nx * s.ny;
  const size_t bytes = plane * (size_t)s.nz;
  if (bytes < kSlabMinBytes || plane > kSlabChunk || s.kz < 1) return p;
  // largest compile-time z length P with m = P - (kz - 1) >= 2 (kz - 1) and
  // at least four slabs
  int lens[64];
  const int cnt = fcamd_static_lengths(lens, 64);
  for (int i = 0; i < cnt && i < 64; ++i) {
    const int P = lens[i], m = P - (s.kz - 1);
    if (m < 8 || m < 2 * (s.kz - 1)) continue;
    if ((s.nz + m - 1) / m < 4) continue;
    if (P > p.P) {
      p.P = P;
      p.m = m;
    }
  }
  if (p.m > 0) {
    // balanced slabs, all transformed at one compile-time length
    p.nslabs = (s.nz + p.m - 1) / p.m;
    p.m = (s.nz + p.nslabs - 1) / p.nslabs;
    p.P = fast_length(p.m + s.kz - 1);
  }
  return p;
}

PaddedProblem slab_problem(const ConvShape& s, int z0, int m, int P) {
  PaddedProblem p{};
  p.padded = s;
  VPad& v = p.vp;
  v.on = 1;
  const int n[3] = {s.nx, s.ny, s.nz};
  for (int d = 0; d < 3; ++d) {
    v.mode[d] = kMapPeriodic;
    v.n[d] = n[d];
    v.m[d] = n[d];
    v.lo[d] = 0;
    v.off[d] = 0;
  }
  v.lo[2] = s.kz - 1 - s.cz;
  v.off[2] = z0;
  v.m[2] = m;
  FC_REQUIRE(P >= m + s.kz - 1, kInternal, "slab transform too short");
  p.padded.nz = P;
  // the window the slab needs; planes beyond it read zero (they may not
  // have been uploaded yet)
  v.hi[0] = s.nx;
  v.hi[1] = s.ny;
  v.hi[2] = m + s.kz - 1;
  v.py = s.ny;
  validate_shape(p.padded);
  return p;
}

size_t slab_workspace_bytes(const ConvShape& s, const SlabPlan& sp) {
  size_t w = 0;
  for (int i = 0; i < sp.nslabs; ++i) {
    const int z0 = i * sp.m, m = std::min(sp.m, s.nz - z0);
    w = std::max(w, workspace_layout(slab_problem(s, z0, m, sp.P).padded).total());
  }
  return w;
}

void host_convolution_slabs(const ConvShape& s, const SlabPlan& sp, const float* im,
                            const float* kernel, float* out, Op op, int dev, float* d_in,
                            float* d_out, float* d_ker, void* ws, hipStream_t compute,
                            void (*host_copy)(void*, const void*, size_t)) {
  const size_t plane = sizeof(float) * (size_t)s.nx * s.ny;
  const int cp = std::max<int>(1, (int)(kSlabChunk / plane));  // planes per chunk
  SlabStreams& ss = slab_streams(dev, plane * cp);
  const int lo = s.kz - 1 - s.cz, nz = s.nz;

  // upload chunks: the tail planes [nz - lo, nz) first, then [0, nz - lo)
  struct Range {
    int z0, n;
  };
  std::vector<Range> up;
  const int head = nz - lo;
  for (int z = head; z < nz; z += cp) up.push_back({z, std::min(cp, nz - z)});
  const size_t ntail = up.size();
  for (int z = 0; z < head; z += cp) up.push_back({z, std::min(cp, head - z)});
  // planes of [0, head) covered after upload chunk i (tail chunks cover none)
  auto covered = [&](size_t i) { return i < ntail ? 0 : up[i].z0 + up[i].n; };

  // download chunks, slab by slab
  struct Down {
    int slab, z0, n;
  };
  std::vector<Down> down;
  for (int k = 0; k < sp.nslabs; ++k) {
    const int z0 = k * sp.m, m = std::min(sp.m, nz - z0);
    for (int z = z0; z < z0 + m; z += cp) down.push_back({k, z, std::min(cp, z0 + m - z)});
  }

  EventPool ev_up, ev_conv, ev_down;
  FC_HIP_CHECK(hipMemcpyAsync(d_ker, kernel, sizeof(float) * (size_t)s.kx * s.ky * s.kz,
                              hipMemcpyHostToDevice, ss.up));
  size_t up_next = 0, down_issue = 0, down_done = 0;
  int slab_next = 0, slab_waited_down = -1;
  const size_t up_total = up.size(), down_total = down.size();
  while (down_done < down_total) {
    bool progress = false;
    // 1. one upload chunk if its pinned buffer is free
    if (up_next < up_total) {
      const int b = (int)(up_next % kUpBufs);
      const bool free_buf = up_next < (size_t)kUpBufs || done(ev_up.get(up_next - kUpBufs));
      if (free_buf) {
        const Range r = up[up_next];
        host_copy(ss.ubuf[b], reinterpret_cast<const char*>(im) + plane * r.z0, plane * r.n);
        FC_HIP_CHECK(hipMemcpyAsync(reinterpret_cast<char*>(d_in) + plane * r.z0, ss.ubuf[b],
                                    plane * r.n, hipMemcpyHostToDevice, ss.up));
        FC_HIP_CHECK(hipEventRecord(ev_up.get(up_next), ss.up));
        ++up_next;
        progress = true;
      }
    }
    // 2. every slab whose input planes have all been enqueued for upload
    while (slab_next < sp.nslabs) {
      const int z0 = slab_next * sp.m, m = std::min(sp.m, nz - z0);
      // planes of [0, head) it reads (its window is z0 - lo .. z0 + m + cz - 1
      // mod nz; the tail planes went first)
      const int need = std::min(z0 + m + s.cz, head);
      if (up_next < ntail || up_next == 0) break;
      size_t last = 0;
      bool ready = false;
      for (size_t i = ntail; i < up_next; ++i)
        if (covered(i) >= need) {
          last = i;
          ready = true;
          break;
        }
      if (!ready) break;
      FC_HIP_CHECK(hipStreamWaitEvent(compute, ev_up.get(last), 0));
      const PaddedProblem p = slab_problem(s, z0, m, sp.P);
      enqueue_convolution_padded(p, d_out + (size_t)z0 * s.nx * s.ny, d_in, d_ker, op, ws,
                                 compute);
      FC_HIP_CHECK(hipEventRecord(ev_conv.get(slab_next), compute));
      ++slab_next;
      progress = true;
    }
    // 3. issue downloads of computed slabs into free pinned buffers
    while (down_issue < down_total && down_issue - down_done < (size_t)kDownBufs &&
           down[down_issue].slab < slab_next) {
      const Down d = down[down_issue];
      if (d.slab != slab_waited_down) {
        FC_HIP_CHECK(hipStreamWaitEvent(ss.down, ev_conv.get(d.slab), 0));
        slab_waited_down = d.slab;
      }
      const int b = (int)(down_issue % kDownBufs);
      FC_HIP_CHECK(hipMemcpyAsync(ss.dbuf[b], reinterpret_cast<const char*>(d_out) + plane * d.z0,
                                  plane * d.n, hipMemcpyDeviceToHost, ss.down));
      FC_HIP_CHECK(hipEventRecord(ev_down.get(down_issue), ss.down));
      ++down_issue;
      progress = true;
    }
    // 4. copy out finished downloads (in order)
    while (down_done < down_issue && done(ev_down.get(down_done))) {
      const Down d = down[down_done];
      host_copy(reinterpret_cast<char*>(out) + plane * d.z0, ss.dbuf[down_done % kDownBufs],
                plane * d.n);
      ++down_done;
      progress = true;
    }
    if (!progress) {
      // nothing could move: wait for the oldest transfer in flight
      if (down_done < down_issue)
        FC_HIP_CHECK(hipEventSynchronize(ev_down.get(down_done)));
      else if (up_next < up_total && up_next >= (size_t)kUpBufs)
        FC_HIP_CHECK(hipEventSynchronize(ev_up.get(up_next - kUpBufs)));
      else if (slab_next > 0)
        FC_HIP_CHECK(hipEventSynchronize(ev_conv.get(slab_next - 1)));
      else
        std::this_thread::yield();
    }
  }
  // the upload and compute streams are idle once every download has landed
  FC_HIP_CHECK(hipStreamSynchronize(ss.up));
  FC_HIP_CHECK(hipStreamSynchronize(compute));
}

}  // namespace fcamd
