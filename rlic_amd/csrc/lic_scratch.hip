// Call scratch blocks (see lic_scratch.h).
#include "lic_scratch.h"

#include <map>
#include <utility>

namespace rlic {

namespace {

struct CachedBlock {
  std::mutex mu;       // held by the call that uses the block
  void* ptr = nullptr;
  size_t bytes = 0;
};

struct BlockCache {
  std::mutex mu;  // guards the map
  std::map<std::pair<int, hipStream_t>, CachedBlock> blocks;
};

BlockCache& cache() {
  static BlockCache* c = new BlockCache;  // never destroyed: blocks outlive static teardown
  return *c;
}

}  // namespace

CallScratch::~CallScratch() {
  if (own_) (void)hipFreeAsync(own_, stream_);
}

hipError_t CallScratch::get(size_t bytes, hipStream_t stream) {
  stream_ = stream;
  if (bytes == 0) bytes = 1;
  int dev = -1;
  if (bytes <= kCachedScratchBytes && hipGetDevice(&dev) == hipSuccess) {
    CachedBlock* b;
    {
      std::lock_guard<std::mutex> g(cache().mu);
      b = &cache().blocks[{dev, stream}];  // std::map nodes do not move
    }
    lock_ = std::unique_lock<std::mutex>(b->mu);
    if (b->bytes < bytes) {
      if (b->ptr) (void)hipFreeAsync(b->ptr, stream);
      b->ptr = nullptr;
      b->bytes = 0;
      const size_t want = (bytes + (size_t{1} << 20) - 1) & ~((size_t{1} << 20) - 1);
      if (hipError_t e = hipMallocAsync(&b->ptr, want, stream); e != hipSuccess) {
        b->ptr = nullptr;
        lock_.unlock();
        return e;
      }
      b->bytes = want;
    }
    ptr_ = b->ptr;
    return hipSuccess;
  }
  (void)hipGetLastError();
  hipError_t e = hipMallocAsync(&own_, bytes, stream);
  if (e != hipSuccess) own_ = nullptr;
  ptr_ = own_;
  return e;
}

hipError_t release_cached_scratch() {
  std::lock_guard<std::mutex> g(cache().mu);
  hipError_t first = hipSuccess;
  int current = -1;
  (void)hipGetDevice(&current);
  for (auto& kv : cache().blocks) {
    CachedBlock& b = kv.second;
    std::lock_guard<std::mutex> busy(b.mu);  // waits for a call that is using it
    if (!b.ptr) continue;
    hipError_t e = hipSetDevice(kv.first.first);
    if (e == hipSuccess) e = hipDeviceSynchronize();  // no call still uses the block
    if (e == hipSuccess) e = hipFreeAsync(b.ptr, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess && first == hipSuccess) first = e;
    b.ptr = nullptr;
    b.bytes = 0;
  }
  if (current >= 0) (void)hipSetDevice(current);
  return first;
}

}  // namespace rlic
