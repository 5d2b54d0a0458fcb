// staging.cpp -- pinned staging of pageable host transfers (staging.hpp).
#include "staging.hpp"

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

namespace fcamd {
namespace {

constexpr size_t kChunk = size_t(8) << 20;  // bytes per staged chunk
constexpr int kBuffers = 4;                  // pinned buffers per device
constexpr size_t kDirect = size_t(4) << 20;  // below this, one plain copy
constexpr int kCopyThreads = 4;              // host copy parallelism

// Fixed pool of worker threads for the host-side copies.  Created on first
// use and never destroyed: at process exit the workers are idle on the
// condition variable and end with the process.
class CopyPool {
 public:
  CopyPool() {
    for (int i = 0; i < kCopyThreads - 1; ++i) workers_.emplace_back([this] { loop(); });
  }
  // copies [src, src+bytes) to dst, split over the workers and the caller
  // (one copy at a time: callers from several devices queue here)
  void copy(void* dst, const void* src, size_t bytes) {
    std::lock_guard<std::mutex> call(call_mu_);
    const size_t part = (bytes / kCopyThreads + 63) / 64 * 64;
    std::vector<std::pair<size_t, size_t>> parts;
    for (size_t off = 0; off < bytes; off += part) parts.push_back({off, std::min(part, bytes - off)});
    if (parts.size() <= 1) {
      std::memcpy(dst, src, bytes);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      pending_ = parts.size() - 1;
      for (size_t i = 1; i < parts.size(); ++i)
        jobs_.push_back([=] {
          std::memcpy(static_cast<char*>(dst) + parts[i].first,
                      static_cast<const char*>(src) + parts[i].first, parts[i].second);
        });
    }
    cv_.notify_all();
    std::memcpy(dst, src, parts[0].second);
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [this] { return pending_ == 0; });
  }

 private:
  void loop() {
    for (;;) {
      std::function<void()> job;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [this] { return !jobs_.empty(); });
        job = std::move(jobs_.back());
        jobs_.pop_back();
      }
      job();
      std::lock_guard<std::mutex> lk(mu_);
      if (--pending_ == 0) done_cv_.notify_all();
    }
  }
  std::mutex call_mu_, mu_;
  std::condition_variable cv_, done_cv_;
  std::vector<std::function<void()>> jobs_;
  size_t pending_ = 0;
  std::vector<std::thread> workers_;
};

CopyPool& copy_pool() {
  static CopyPool* p = new CopyPool();  // intentionally never destroyed
  return *p;
}

// one caller at a time per device (the pinned buffers are reused)
struct Stager {
  std::mutex mu;
  void* buf[kBuffers] = {};
  hipEvent_t done[kBuffers] = {};
};

std::mutex g_stagers_mu;
std::map<int, Stager*> g_stagers;  // never freed: pinned memory lives with the process

Stager& stager(int dev) {
  std::lock_guard<std::mutex> lk(g_stagers_mu);
  Stager*& s = g_stagers[dev];
  if (!s) {
    auto* n = new Stager();
    for (int i = 0; i < kBuffers; ++i) {
      FC_HIP_CHECK(hipHostMalloc(&n->buf[i], kChunk, hipHostMallocDefault));
      FC_HIP_CHECK(hipEventCreateWithFlags(&n->done[i], hipEventDisableTiming));
    }
    s = n;
  }
  return *s;
}

}  // namespace

void host_copy_parallel(void* dst, const void* src, size_t bytes) { copy_pool().copy(dst, src, bytes); }

void staged_upload(int dev, void* d_dst, const void* h_src, size_t bytes, hipStream_t st) {
  if (bytes < kDirect) {
    FC_HIP_CHECK(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, st));
    return;
  }
  Stager& s = stager(dev);
  std::lock_guard<std::mutex> lk(s.mu);
  const size_t n = (bytes + kChunk - 1) / kChunk;
  for (size_t i = 0; i < n; ++i) {
    const int b = (int)(i % kBuffers);
    const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
    FC_HIP_CHECK(hipEventSynchronize(s.done[b]));  // previous DMA from this buffer
    copy_pool().copy(s.buf[b], static_cast<const char*>(h_src) + off, len);
    FC_HIP_CHECK(hipMemcpyAsync(static_cast<char*>(d_dst) + off, s.buf[b], len,
                                hipMemcpyHostToDevice, st));
    FC_HIP_CHECK(hipEventRecord(s.done[b], st));
  }
}

void staged_download(int dev, void* h_dst, const void* d_src, size_t bytes, hipStream_t st) {
  if (bytes < kDirect) {
    FC_HIP_CHECK(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, st));
    FC_HIP_CHECK(hipStreamSynchronize(st));
    return;
  }
  Stager& s = stager(dev);
  std::lock_guard<std::mutex> lk(s.mu);
  const size_t n = (bytes + kChunk - 1) / kChunk;
  auto issue = [&](size_t i) {
    const int b = (int)(i % kBuffers);
    const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
    FC_HIP_CHECK(hipMemcpyAsync(s.buf[b], static_cast<const char*>(d_src) + off, len,
                                hipMemcpyDeviceToHost, st));
    FC_HIP_CHECK(hipEventRecord(s.done[b], st));
  };
  for (size_t i = 0; i < n && i < (size_t)kBuffers; ++i) issue(i);
  for (size_t i = 0; i < n; ++i) {
    const int b = (int)(i % kBuffers);
    const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
    FC_HIP_CHECK(hipEventSynchronize(s.done[b]));
    copy_pool().copy(static_cast<char*>(h_dst) + off, s.buf[b], len);
    if (i + kBuffers < n) issue(i + kBuffers);
  }
}

}  // namespace fcamd
