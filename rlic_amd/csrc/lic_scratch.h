// Device scratch of one library call, from one block.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include <mutex>

namespace rlic {

// Blocks up to this size are cached per (device, stream) and reused.
constexpr size_t kCachedScratchBytes = size_t{256} << 20;

// Scratch for one call's work on `stream`.  A block of at most
// kCachedScratchBytes comes from a grow-only cache kept per (device,
// stream): the next call on that stream reuses it, which is safe because
// its work is ordered after this call's on the stream.  The cache entry
// stays locked while the CallScratch lives, so two host threads enqueueing
// on one stream take turns.  Larger blocks are allocated from the
// stream-ordered pool and released on the stream when the CallScratch ends.
// Cached blocks are never freed (they live until the process ends).
class CallScratch {
 public:
  CallScratch() = default;
  CallScratch(const CallScratch&) = delete;
  CallScratch& operator=(const CallScratch&) = delete;
  ~CallScratch();
  hipError_t get(size_t bytes, hipStream_t stream);
  void* ptr() const { return ptr_; }

 private:
  void* ptr_ = nullptr;
  void* own_ = nullptr;          // per-call block (above the cache limit)
  hipStream_t stream_ = nullptr;
  std::unique_lock<std::mutex> lock_;
};

// Frees every cached block (after synchronizing the device the blocks
// belong to); the next calls allocate again.
hipError_t release_cached_scratch();

}  // namespace rlic
