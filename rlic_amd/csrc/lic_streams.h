// Side streams: independent launches spread over a few streams so that
// launches too small to fill the GPU (and their tails) overlap.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

namespace rlic {

// Side streams of the current device: default priority, non-blocking,
// created on first use and kept until the process ends (shared by every
// caller on the device; the work of different callers is ordered by their
// own events).  start() orders the side streams after the work enqueued
// so far on the caller's stream; join() makes the caller's stream wait for
// everything enqueued on them since.  The destructor waits on the host for
// work that was never joined (error paths).
class SideWork {
 public:
  explicit SideWork(hipStream_t caller) : caller_(caller) {}
  SideWork(const SideWork&) = delete;
  SideWork& operator=(const SideWork&) = delete;
  ~SideWork() { finish(); }

  std::vector<hipStream_t> streams;  // empty: not started (use the caller's stream)

  bool start(int dev, int k) {
    static std::mutex mu;
    static std::map<int, std::vector<hipStream_t>> pool;
    {
      std::lock_guard<std::mutex> lock(mu);
      std::vector<hipStream_t>& p = pool[dev];
      while (static_cast<int>(p.size()) < k) {
        hipStream_t s = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
          (void)hipGetLastError();
          return false;
        }
        p.push_back(s);
      }
      streams.assign(p.begin(), p.begin() + k);
    }
    hipEvent_t ready = nullptr;
    bool ok = hipEventCreateWithFlags(&ready, hipEventDisableTiming) == hipSuccess &&
              hipEventRecord(ready, caller_) == hipSuccess;
    for (hipStream_t s : streams) ok = ok && hipStreamWaitEvent(s, ready, 0) == hipSuccess;
    if (ready) (void)hipEventDestroy(ready);  // released once the waits are satisfied
    if (!ok) {
      (void)hipGetLastError();
      finish();
      return false;
    }
    return true;
  }

  // The caller's stream waits for the side streams' work; the first error.
  hipError_t join() {
    hipError_t first = hipSuccess;
    for (hipStream_t s : streams) {
      hipEvent_t done = nullptr;
      hipError_t e = hipEventCreateWithFlags(&done, hipEventDisableTiming);
      if (e == hipSuccess) e = hipEventRecord(done, s);
      if (e == hipSuccess) e = hipStreamWaitEvent(caller_, done, 0);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        e = hipStreamSynchronize(s);  // no event: wait on the host instead
      }
      if (done) (void)hipEventDestroy(done);
      if (e != hipSuccess && first == hipSuccess) first = e;
    }
    streams.clear();
    return first;
  }

  // Waits on the host for the side streams' work (error paths).
  void finish() {
    if (streams.empty()) return;
    for (hipStream_t s : streams) (void)hipStreamSynchronize(s);
    (void)hipGetLastError();
    streams.clear();
  }

  // Stream of each of `costs.size()` launches over k streams: largest cost
  // first, each to the least loaded stream.
  static std::vector<int> assign(const std::vector<int64_t>& costs, int k) {
    std::vector<size_t> order(costs.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return costs[a] > costs[b]; });
    std::vector<int64_t> load(static_cast<size_t>(std::max(k, 1)), 0);
    std::vector<int> lane(costs.size(), 0);
    for (size_t i : order) {
      const size_t s =
          static_cast<size_t>(std::min_element(load.begin(), load.end()) - load.begin());
      lane[i] = static_cast<int>(s);
      load[s] += costs[i];
    }
    return lane;
  }

 private:
  hipStream_t caller_;
};

}  // namespace rlic
