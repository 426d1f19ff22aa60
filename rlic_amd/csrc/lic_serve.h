// Serving core (lic_serve.hip): fields resident in device memory, advanced
// one pass per step, many fields per step in one persistent launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "lic_device.h"

namespace rlic {

// One replay pass (src -> dst along the move record `codes`) of a field as
// a single-field launch with the default replay plan (LDS box, hybrid or L2
// gathers; no tuning).  `kernel`: the first weight of a copy padded with
// kernel_pad() zeros on each side.  Defined in lic_api.hip.
template <typename T>
hipError_t replay_pass(const T* src, T* dst, const T* kernel, int32_t klen, const FieldGeom& g,
                       int profile, const uint4* codes, hipStream_t stream);

// Device state of one admitted field, in one allocation:
// [padded weights | cells | buffer 0 | buffer 1 | move record].  The weights
// are kept with kernel_pad() zeros on each side (the single-field replay
// kernels read them so); `weights` points at the first weight.
template <typename T>
struct StepWork;
class SideWork;

template <typename T>
struct Resident {
  FieldGeom g{};
  int32_t klen = 0, uv_mode = 0;
  void* block = nullptr;
  size_t bytes = 0;        // budget charge (field_bytes)
  size_t block_bytes = 0;  // allocated: bytes rounded up to a size class
  T* weights = nullptr;
  Cell<T>* cells = nullptr;
  T* buf[2] = {nullptr, nullptr};
  uint4* codes = nullptr;  // nullptr: the field keeps no move record
  int64_t done = 0;        // passes computed
  int64_t target = 0;      // passes requested so far
  int cur = -1;            // buffer holding the latest output (-1: none yet)
  bool recorded = false;   // `codes` holds the move record
};

class Server {
 public:
  Server(int64_t budget_bytes, int profile, hipStream_t stream)
      : budget_(budget_bytes), profile_(profile), stream_(stream) {}
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  // Bytes one field occupies while resident.
  template <typename T>
  static size_t field_bytes(int64_t ny, int64_t nx, int64_t klen, bool record);

  // Copies the kernel weights, packs (u, v, texture) and registers the field
  // for `iterations` (> 0) passes.  Returns RLIC_ERR_OUT_OF_MEMORY when the
  // field does not fit the remaining budget (nothing is allocated then).
  // `record`: keep the move record, for more than one pass or follow-ups.
  template <typename T>
  int admit(const T* tex, int64_t tsy, int64_t tsx, const T* u, int64_t usy, int64_t usx,
            const T* v, int64_t vsy, int64_t vsx, int64_t ny, int64_t nx, const T* kernel,
            int64_t klen, int uv_mode, int xl, int xr, int yl, int yr, int64_t iterations,
            bool record, int64_t* handle, std::string& err);

  // One pass of every field with passes left, in one launch per dtype;
  // appends the handles whose requested passes are all done.  Synchronous.
  int step(std::vector<int64_t>& finished, std::string& err);

  // Requests `more` passes of a resident field whose passes are all done
  // (it must keep its move record, unless it has not run yet).
  int extend(int64_t handle, int64_t more, std::string& err);

  // Copies the field's latest output (ny x nx, C order) to `out`, device
  // or host memory.  Synchronous.
  template <typename T>
  int read(int64_t handle, T* out, std::string& err);

  int release(int64_t handle, std::string& err);

  void usage(int64_t* used_bytes, int64_t* resident, int64_t* active);
  hipStream_t stream() const { return stream_; }

 private:
  template <typename T>
  std::map<int64_t, Resident<T>>& table();
  template <typename T>
  void plan_dtype(StepWork<T>& w, std::vector<int64_t>& finished);
  template <typename T>
  int launch_singles(StepWork<T>& w, const SideWork& side, const int* lane, std::string& err);
  template <typename T>
  void commit_dtype(StepWork<T>& w, std::vector<int64_t>& finished);

  // Resident blocks come from a cache of released ones (same size class)
  // before the stream-ordered allocator: with blocks of many sizes coming
  // and going, hipMallocAsync and hipFreeAsync cost 80 and 220 us a call.
  // Cached blocks count against the budget as free memory held:
  // used_ + cached_bytes_ <= budget_ (the cache is trimmed to keep it).
  static size_t size_class(size_t bytes);
  hipError_t take_block(size_t charge, size_t cls, void** block);
  void give_block(void* block, size_t cls);
  void trim_cache(size_t keep_room);

  std::mutex mu_;
  int64_t budget_;
  int64_t used_ = 0;
  int profile_;
  hipStream_t stream_;
  int64_t next_handle_ = 1;
  std::map<int64_t, Resident<float>> f32_;
  std::map<int64_t, Resident<double>> f64_;
  std::multimap<size_t, void*> cache_;  // size class -> released block
  size_t cached_bytes_ = 0;
};

}  // namespace rlic
