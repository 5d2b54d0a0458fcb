// slabs.cpp -- host-pointer convolution streamed through the device in
// z-slabs (overlap-save), holding only a ring of slab buffers.
//
// The output planes are cut into slabs of m planes.  Slab [z0, z0 + m) needs
// the input window of planes (z0 - lo + p) mod nz, p < W = m + kz - 1
// (lo = kz - 1 - kz/2): its circular convolution, transformed at a z length
// P >= W with the window's planes beyond W reading zero, equals the circular
// convolution of the whole volume on the slab's m planes.  x and y are not
// padded.  So the device never holds a whole volume:
//   * a ring of kRingBuffers window buffers of W planes: slab k's window
//     goes into buffer k mod kRingBuffers.  Consecutive windows overlap by
//     kz - 1 planes (the halo), which are copied on the device from the
//     previous slab's buffer; only the m new planes cross PCIe.
//   * the slab's output is written into planes [0, m) of its own buffer
//     (the x-C2R runs after the x-R2C has read the window), leaving the
//     halo planes [m, W) intact for the next slab, and is downloaded from
//     there.
//   * the last slab's window wraps to planes [0, kz/2); those are copied
//     into a small wrap buffer from slab 0's window before slab 0's output
//     overwrites them, because an in-place call writes the first slab's
//     output over the host planes before the last slab is uploaded.
// Uploads, slab convolutions and downloads run on three streams driven by
// one host loop: while slab k computes, slab k+1 uploads and slab k-1
// downloads.  All transfers go through pinned buffers.
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

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

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

size_t plane_bytes(const ConvShape& s) { return sizeof(float) * (size_t)s.nx * s.ny; }

// buffer sizes of a slab plan with output slabs of m planes at length P
void size_plan(const ConvShape& s, SlabPlan& p) {
  const size_t W = (size_t)p.m + s.kz - 1;
  p.ring_bytes = kRingBuffers * align256(W * plane_bytes(s));
  p.wrap_bytes = align256(std::max<size_t>(1, (size_t)s.cz) * plane_bytes(s));
  p.ws_bytes = workspace_layout(slab_window_problem(s, p.m, p.P).padded).total();
}

}  // namespace

size_t single_shot_device_bytes(const ConvShape& s) {
  return align256(sizeof(float) * (size_t)s.voxels()) + workspace_layout(s).total() +
         align256(sizeof(float) * (size_t)s.kx * s.ky * s.kz);
}

SlabPlan slab_plan(const ConvShape& s, size_t budget, bool force) {
  SlabPlan best{};
  const size_t bytes = plane_bytes(s) * (size_t)s.nz;
  if ((!force && bytes < kSlabMinBytes) || s.kz < 1) return best;
  const int min_slabs = force ? 2 : 4;
  int lens[128];
  const int cnt = fcamd_static_lengths(lens, 128);
  for (int i = 0; i < cnt && i < 128; ++i) {
    const int m0 = lens[i] - (s.kz - 1);
    if (m0 < 8 || m0 < 2 * (s.kz - 1)) continue;
    const int n = (s.nz + m0 - 1) / m0;
    if (n < min_slabs) continue;
    // balanced slabs, all transformed at one length
    SlabPlan c{};
    c.nslabs = n;
    c.m = (s.nz + n - 1) / n;
    c.P = fast_length(c.m + s.kz - 1);
    // the first window must not wrap onto itself
    if (c.m + s.kz - 1 > s.nz) continue;
    size_plan(s, c);
    if (c.device_bytes() > budget) continue;
    if (c.m > best.m) best = c;
  }
  return best;
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
  // the window the slab needs; planes beyond it read zero
  v.hi[0] = s.nx;
  v.hi[1] = s.ny;
  v.hi[2] = m + s.kz - 1;
  v.py = s.ny;
  validate_shape(p.padded);
  return p;
}

PaddedProblem slab_window_problem(const ConvShape& s, int m, int P) {
  const int lo = s.kz - 1 - s.cz, W = m + s.kz - 1;
  PaddedProblem p = slab_problem(s, 0, m, P);
  // padded position p reads buffer plane p - lo + off = p (p < W), and the
  // cropped output plane p - lo goes to buffer plane p - lo
  p.vp.n[2] = W;
  p.vp.off[2] = lo;
  return p;
}

void host_convolution_slabs(const ConvShape& s, const SlabPlan& sp, const float* im,
                            const float* kernel, float* out, Op op, int dev, float* ring,
                            float* wrap, float* d_ker, void* ws, hipStream_t compute,
                            void (*host_copy)(void*, const void*, size_t)) {
  const size_t plane = plane_bytes(s);
  const int cp = std::max<int>(1, (int)(kSlabChunk / plane));  // planes per chunk
  SlabStreams& ss = slab_streams(dev, plane * cp);
  const int nz = s.nz, cz = s.cz, lo = s.kz - 1 - s.cz, halo = s.kz - 1;
  const int NB = kRingBuffers;
  const size_t stride = sp.ring_bytes / NB;  // bytes per ring buffer
  FC_REQUIRE(stride >= ((size_t)sp.m + halo) * plane, kInternal, "ring buffer too small");
  auto buf = [&](int k) { return reinterpret_cast<char*>(ring) + (size_t)(k % NB) * stride; };

  struct Slab {
    int z0, m;
    int first_up = -1, last_up = -1;      // upload chunk range
    int first_down = -1, last_down = -1;  // download chunk range
  };
  std::vector<Slab> slabs(sp.nslabs);
  for (int k = 0; k < sp.nslabs; ++k) {
    slabs[k].z0 = k * sp.m;
    slabs[k].m = std::min(sp.m, nz - slabs[k].z0);
    FC_REQUIRE(slabs[k].m >= 1, kInternal, "empty slab");
  }

  // upload chunks: host planes [src, src + n) into buffer planes [dst, dst + n)
  struct Up {
    int slab, src, dst, n;
  };
  std::vector<Up> up;
  auto add_up = [&](int k, int src, int dst, int n) {
    for (int i = 0; i < n; i += cp) {
      if (slabs[k].first_up < 0) slabs[k].first_up = (int)up.size();
      slabs[k].last_up = (int)up.size();
      up.push_back({k, src + i, dst + i, std::min(cp, n - i)});
    }
  };
  for (int k = 0; k < sp.nslabs; ++k) {
    const Slab& sl = slabs[k];
    if (k == 0) {
      // the whole window: the tail planes [nz - lo, nz), then [0, m + cz)
      add_up(0, nz - lo, 0, lo);
      add_up(0, 0, lo, sl.m + cz);
    } else {
      // the m new planes (window positions halo..W-1, sources z0 + cz ..),
      // those below nz from the host; the rest wrap to the saved planes
      const int s0 = sl.z0 + cz, s1 = std::min(sl.z0 + sl.m + cz, nz);
      if (s1 > s0) add_up(k, s0, halo, s1 - s0);
    }
  }
  // download chunks: buffer planes [0, m) of each slab to host [z0, z0 + m)
  struct Down {
    int slab, z, n;
  };
  std::vector<Down> down;
  for (int k = 0; k < sp.nslabs; ++k) {
    Slab& sl = slabs[k];
    for (int i = 0; i < sl.m; i += cp) {
      if (sl.first_down < 0) sl.first_down = (int)down.size();
      sl.last_down = (int)down.size();
      down.push_back({k, i, std::min(cp, sl.m - i)});
    }
  }

  EventPool ev_up, ev_conv, ev_down;
  FC_HIP_CHECK(hipMemcpyAsync(d_ker, kernel, sizeof(float) * (size_t)s.kx * s.ky * s.kz,
                              hipMemcpyHostToDevice, ss.up));
  const size_t up_total = up.size(), down_total = down.size();
  size_t up_next = 0, down_issue = 0, down_done = 0;
  int slab_next = 0, slab_waited_down = -1;
  // downloads of slab k all enqueued (its last-chunk event recorded)
  auto downs_issued = [&](int k) { return (int)down_issue > slabs[k].last_down; };
  while (down_done < down_total) {
    bool progress = false;
    // 1. one upload chunk, once its pinned buffer is free and its ring
    //    buffer's previous slab no longer needs it
    if (up_next < up_total) {
      const Up& u = up[up_next];
      const int k = u.slab;
      const bool first = (int)up_next == slabs[k].first_up;
      // buffer k mod NB last held slab k - NB (downloaded from it) and fed
      // slab k - NB + 1 its halo
      const bool ring_free = k < NB || (downs_issued(k - NB) && slab_next > k - NB + 1);
      const bool pinned_free =
          up_next < (size_t)kUpBufs || done(ev_up.get(up_next - kUpBufs));
      if (ring_free && pinned_free) {
        if (first && k >= NB) {
          FC_HIP_CHECK(hipStreamWaitEvent(ss.up, ev_down.get(slabs[k - NB].last_down), 0));
          FC_HIP_CHECK(hipStreamWaitEvent(ss.up, ev_conv.get(k - NB + 1), 0));
        }
        const int b = (int)(up_next % kUpBufs);
        host_copy(ss.ubuf[b], reinterpret_cast<const char*>(im) + plane * u.src, plane * u.n);
        FC_HIP_CHECK(hipMemcpyAsync(buf(k) + plane * u.dst, ss.ubuf[b], plane * u.n,
                                    hipMemcpyHostToDevice, ss.up));
        FC_HIP_CHECK(hipEventRecord(ev_up.get(up_next), ss.up));
        ++up_next;
        progress = true;
      }
    }
    // 2. the next slab's convolution once its uploads are enqueued
    while (slab_next < sp.nslabs) {
      const int k = slab_next;
      const Slab& sl = slabs[k];
      if (sl.last_up >= 0 && (int)up_next <= sl.last_up) break;
      if (k >= NB && !downs_issued(k - NB)) break;
      if (sl.last_up >= 0) FC_HIP_CHECK(hipStreamWaitEvent(compute, ev_up.get(sl.last_up), 0));
      if (k >= NB)  // the halo copy below overwrites planes slab k - NB is downloaded from
        FC_HIP_CHECK(hipStreamWaitEvent(compute, ev_down.get(slabs[k - NB].last_down), 0));
      if (k == 0 && cz > 0)  // keep the planes the last window wraps to
        FC_HIP_CHECK(hipMemcpyAsync(wrap, buf(0) + plane * lo, plane * cz,
                                    hipMemcpyDeviceToDevice, compute));
      if (k > 0 && halo > 0) {
        const Slab& pr = slabs[k - 1];
        FC_HIP_CHECK(hipMemcpyAsync(buf(k), buf(k - 1) + plane * pr.m, plane * halo,
                                    hipMemcpyDeviceToDevice, compute));
      }
      if (k > 0) {
        // window sources at or past nz wrap to the saved planes [0, cz)
        const int s_lo = std::max(sl.z0 + cz, nz), s_hi = sl.z0 + sl.m + cz;
        if (s_hi > s_lo) {
          const int p0 = s_lo - (sl.z0 - lo);
          FC_HIP_CHECK(hipMemcpyAsync(buf(k) + plane * p0,
                                      reinterpret_cast<const char*>(wrap) + plane * (s_lo - nz),
                                      plane * (s_hi - s_lo), hipMemcpyDeviceToDevice, compute));
        }
      }
      const PaddedProblem p = slab_window_problem(s, sl.m, sp.P);
      float* window = reinterpret_cast<float*>(buf(k));
      enqueue_convolution_padded(p, window, window, d_ker, op, ws, compute);
      FC_HIP_CHECK(hipEventRecord(ev_conv.get(k), compute));
      ++slab_next;
      progress = true;
    }
    // 3. downloads of computed slabs into free pinned buffers
    while (down_issue < down_total && down_issue - down_done < (size_t)kDownBufs &&
           down[down_issue].slab < slab_next) {
      const Down d = down[down_issue];
      if (d.slab != slab_waited_down) {
        FC_HIP_CHECK(hipStreamWaitEvent(ss.down, ev_conv.get(d.slab), 0));
        slab_waited_down = d.slab;
      }
      const int b = (int)(down_issue % kDownBufs);
      FC_HIP_CHECK(hipMemcpyAsync(ss.dbuf[b], buf(d.slab) + plane * d.z, plane * d.n,
                                  hipMemcpyDeviceToHost, ss.down));
      FC_HIP_CHECK(hipEventRecord(ev_down.get(down_issue), ss.down));
      ++down_issue;
      progress = true;
    }
    // 4. copy out finished downloads (in order)
    while (down_done < down_issue && done(ev_down.get(down_done))) {
      const Down d = down[down_done];
      host_copy(reinterpret_cast<char*>(out) + plane * ((size_t)slabs[d.slab].z0 + d.z),
                ss.dbuf[down_done % kDownBufs], plane * d.n);
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
