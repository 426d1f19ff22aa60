// Serving core: resident fields advanced one pass per step.
//
// A request is one `rlic.convolve` call (src/lib.rs:408-443).  Admission
// copies its kernel weights and packs (u, v, texture) into device memory
// that stays resident until the field is released; each step then runs one
// pass of every field that has passes left, all fields of a dtype in one
// persistent launch over square regions (launch_queue without waits: the
// step boundary orders the passes), except the replay passes of fields of
// at least 2^18 pixels, which run as single-field launches (LDS box or
// hybrid replay, as a single convolve call would).  A field's first pass traces and, when
// it keeps a move record, records the moves; later passes (and follow-up
// passes requested after the field finished) replay them.  The two output
// buffers alternate, so a field's passes give the same bits as one
// convolve call with the total number of iterations, whatever else shares
// the step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "../../include/rlic_hip.h"
#include "lic_batch.h"
#include "lic_launch.h"
#include "lic_serve.h"
#include "lic_streams.h"

namespace rlic {

static size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

template <>
std::map<int64_t, Resident<float>>& Server::table<float>() { return f32_; }
template <>
std::map<int64_t, Resident<double>>& Server::table<double>() { return f64_; }

Server::~Server() {
  for (auto& kv : f32_) (void)hipFreeAsync(kv.second.block, stream_);
  for (auto& kv : f64_) (void)hipFreeAsync(kv.second.block, stream_);
  for (auto& kv : cache_) (void)hipFreeAsync(kv.second, stream_);
  (void)hipStreamSynchronize(stream_);
}

// Eight classes per power of two above 1 MiB (at most 12.5% over the
// request), 64 KiB steps below.
size_t Server::size_class(size_t bytes) {
  size_t step = size_t{64} << 10;
  if (bytes > (size_t{1} << 20)) {
    int top = 63 - __builtin_clzll(static_cast<unsigned long long>(bytes));
    step = size_t{1} << (top - 3);
  }
  return (bytes + step - 1) / step * step;
}

// Frees cached blocks, largest first, until used_ + cached_bytes_ +
// keep_room fits the budget (or the cache is empty).  Stream-ordered: a
// freed block's last use was on stream_, or on a side stream the step
// joined before returning.
void Server::trim_cache(size_t keep_room) {
  while (!cache_.empty() &&
         static_cast<int64_t>(cached_bytes_ + keep_room) + used_ > budget_) {
    auto last = std::prev(cache_.end());
    (void)hipFreeAsync(last->second, stream_);
    cached_bytes_ -= last->first;
    cache_.erase(last);
  }
}

hipError_t Server::take_block(size_t charge, size_t cls, void** block) {
  auto it = cache_.find(cls);
  if (it != cache_.end()) {  // reused in stream order after its last use
    *block = it->second;
    cached_bytes_ -= cls;
    cache_.erase(it);
    return hipSuccess;
  }
  trim_cache(charge);
  return hipMallocAsync(block, cls, stream_);
}

void Server::give_block(void* block, size_t cls) {
  cache_.emplace(cls, block);
  cached_bytes_ += cls;
  trim_cache(0);
}

// Replay passes of fields this large run as single-field launches.
constexpr int64_t kSingleReplayMinPixels = int64_t{1} << 18;

template <typename T>
size_t Server::field_bytes(int64_t ny, int64_t nx, int64_t klen, bool record) {
  const size_t n = static_cast<size_t>(ny * nx);
  size_t b = align256(sizeof(T) * static_cast<size_t>(std::max<int64_t>(klen, 0) + 2 * kernel_pad()));
  b += align256(sizeof(Cell<T>) * static_cast<size_t>(guarded_cells(ny, nx)));
  b += 2 * align256(sizeof(T) * n);
  if (record) b += align256(sizeof(uint4) * static_cast<size_t>(code_chunks(static_cast<int32_t>(klen))) * n);
  return b;
}

template <typename T>
int Server::admit(const T* tex, int64_t tsy, int64_t tsx, const T* u, int64_t usy, int64_t usx,
                  const T* v, int64_t vsy, int64_t vsx, int64_t ny, int64_t nx, const T* kernel,
                  int64_t klen, int uv_mode, int xl, int xr, int yl, int yr, int64_t iterations,
                  bool record, int64_t* handle, std::string& err) {
  std::lock_guard<std::mutex> lock(mu_);
  const size_t bytes = field_bytes<T>(ny, nx, klen, record);
  if (used_ + static_cast<int64_t>(bytes) > budget_) {
    err = "the field needs " + std::to_string(bytes) + " bytes; " +
          std::to_string(budget_ - used_) + " of the serving budget are free";
    return RLIC_ERR_OUT_OF_MEMORY;
  }
  Resident<T> r;
  r.g = FieldGeom{static_cast<int32_t>(ny), static_cast<int32_t>(nx), xl, xr, yl, yr};
  r.klen = static_cast<int32_t>(klen);
  r.uv_mode = uv_mode;
  r.bytes = bytes;
  r.block_bytes = size_class(bytes);
  if (hipError_t e = take_block(bytes, r.block_bytes, &r.block); e != hipSuccess) {
    (void)hipGetLastError();
    err = std::string("allocating resident field: ") + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? RLIC_ERR_OUT_OF_MEMORY : RLIC_ERR_HIP;
  }
  const size_t n = static_cast<size_t>(ny * nx);
  char* p = static_cast<char*>(r.block);
  const int32_t pad = kernel_pad();
  T* padded = reinterpret_cast<T*>(p);
  r.weights = padded + pad;
  p += align256(sizeof(T) * static_cast<size_t>(std::max<int64_t>(klen, 0) + 2 * pad));
  r.cells = reinterpret_cast<Cell<T>*>(p);
  p += align256(sizeof(Cell<T>) * static_cast<size_t>(guarded_cells(ny, nx)));
  r.buf[0] = reinterpret_cast<T*>(p);
  p += align256(sizeof(T) * n);
  r.buf[1] = reinterpret_cast<T*>(p);
  p += align256(sizeof(T) * n);
  r.codes = record ? reinterpret_cast<uint4*>(p) : nullptr;
  r.target = iterations;
  // zeros on both sides of the weights (no graph captures this stream)
  hipError_t e = hipMemsetAsync(padded, 0, sizeof(T) * pad, stream_);
  if (e == hipSuccess)
    e = hipMemsetAsync(r.weights + klen, 0, sizeof(T) * pad, stream_);
  if (e == hipSuccess && klen > 0 && n > 0)  // an empty field's kernel is never read
    e = hipMemcpyAsync(r.weights, kernel, sizeof(T) * static_cast<size_t>(klen),
                       hipMemcpyDefault, stream_);
  if (e == hipSuccess && n > 0)
    e = launch_pack<T>(tex, tsy, tsx, u, usy, usx, v, vsy, vsx, r.g, r.cells, stream_);
  if (e != hipSuccess) {
    (void)hipFreeAsync(r.block, stream_);  // not cached: the failed copy may have left it in use
    err = std::string("admitting field: ") + hipGetErrorString(e);
    return RLIC_ERR_HIP;
  }
  used_ += static_cast<int64_t>(bytes);
  *handle = next_handle_++;
  table<T>().emplace(*handle, r);
  return RLIC_OK;
}

// One dtype's share of a step: the fields of the persistent launch and the
// fields whose replay pass runs as a single-field launch.
template <typename T>
struct StepWork {
  std::vector<BatchField<T>> fields;
  std::vector<const T*> ksrc;
  std::vector<Segment> segs;
  std::vector<Resident<T>*> stepped, singles;
  std::vector<int64_t> handles, single_handles;
};

template <typename T>
void Server::plan_dtype(StepWork<T>& w, std::vector<int64_t>& finished) {
  for (auto& kv : table<T>()) {
    Resident<T>& r = kv.second;
    if (r.done >= r.target) continue;
    if (static_cast<int64_t>(r.g.ny) * r.g.nx == 0) {  // nothing to compute
      r.done = r.target;
      finished.push_back(kv.first);
      continue;
    }
    if (r.done > 0 && r.recorded &&
        static_cast<int64_t>(r.g.ny) * r.g.nx >= kSingleReplayMinPixels) {
      w.singles.push_back(&r);
      w.single_handles.push_back(kv.first);
      continue;
    }
    BatchField<T> f;
    std::memset(&f, 0, sizeof(f));
    f.cells = r.cells;
    f.codes = r.codes;
    f.g = r.g;
    f.klen = r.klen;
    f.uv_mode = r.uv_mode;
    f.iterations = static_cast<int32_t>(std::min<int64_t>(r.target - r.done, INT32_MAX));
    Segment sg{};
    sg.field = static_cast<int32_t>(w.fields.size());
    sg.prev = -1;
    const int out_buf = r.cur < 0 ? 0 : 1 - r.cur;
    sg.dst = r.buf[out_buf];
    if (r.done == 0) {
      sg.pass = 0;
      sg.record = r.codes != nullptr ? 1 : 0;
      sg.src = nullptr;
    } else {
      sg.pass = 1;
      sg.record = 0;
      sg.src = r.buf[r.cur];
    }
    w.fields.push_back(f);
    w.ksrc.push_back(r.weights);
    w.segs.push_back(sg);
    w.stepped.push_back(&r);
    w.handles.push_back(kv.first);
  }
}

template <typename T>
int Server::launch_singles(StepWork<T>& w, const SideWork& side, const int* lane,
                           std::string& err) {
  for (size_t k = 0; k < w.singles.size(); ++k) {
    Resident<T>* r = w.singles[k];
    const hipStream_t s = side.streams.empty() ? stream_ : side.streams[lane[k]];
    if (hipError_t e = replay_pass<T>(r->buf[r->cur], r->buf[1 - r->cur], r->weights, r->klen,
                                      r->g, profile_, r->codes, s);
        e != hipSuccess) {
      (void)hipGetLastError();
      err = std::string("replay pass: ") + hipGetErrorString(e);
      return RLIC_ERR_HIP;
    }
  }
  return RLIC_OK;
}

template <typename T>
void Server::commit_dtype(StepWork<T>& w, std::vector<int64_t>& finished) {
  for (size_t k = 0; k < w.stepped.size(); ++k) {
    Resident<T>& r = *w.stepped[k];
    if (r.done == 0 && r.codes != nullptr) r.recorded = true;
    r.cur = r.cur < 0 ? 0 : 1 - r.cur;
    ++r.done;
    if (r.done == r.target) finished.push_back(w.handles[k]);
  }
  for (size_t k = 0; k < w.singles.size(); ++k) {
    Resident<T>& r = *w.singles[k];
    r.cur = 1 - r.cur;
    ++r.done;
    if (r.done == r.target) finished.push_back(w.single_handles[k]);
  }
}

// Both dtypes' work of a step is enqueued before anything is waited for:
// the single-field replays of either dtype go to up to RLIC_BATCH_STREAMS
// (default 4) side streams, largest first to the least loaded one (a
// float64 pixel-step weighs twice a float32 one), where they overlap with
// each other and with the two persistent launches on the server's stream.
int Server::step(std::vector<int64_t>& finished, std::string& err) {
  std::lock_guard<std::mutex> lock(mu_);
  StepWork<float> w32;
  StepWork<double> w64;
  plan_dtype(w32, finished);
  plan_dtype(w64, finished);
  const size_t n_single = w32.singles.size() + w64.singles.size();
  SideWork side(stream_);
  std::vector<int> lane(n_single, 0);
  if (n_single > 1) {
    const char* env = std::getenv("RLIC_BATCH_STREAMS");
    const int want = std::min<int>(std::min(env && *env ? std::atoi(env) : 4, 16),
                                   static_cast<int>(n_single));
    int dev = 0;
    if (want > 1 && hipGetDevice(&dev) == hipSuccess && side.start(dev, want)) {
      std::vector<int64_t> cost;
      for (Resident<float>* r : w32.singles)
        cost.push_back(static_cast<int64_t>(r->g.ny) * r->g.nx * r->klen);
      for (Resident<double>* r : w64.singles)
        cost.push_back(2 * static_cast<int64_t>(r->g.ny) * r->g.nx * r->klen);
      lane = SideWork::assign(cost, want);
    }
    (void)hipGetLastError();
  }
  int rc = launch_singles(w32, side, lane.data(), err);
  if (rc == RLIC_OK) rc = launch_singles(w64, side, lane.data() + w32.singles.size(), err);
  if (rc == RLIC_OK && !w32.fields.empty())
    rc = launch_queue<float>(w32.fields, w32.ksrc, w32.segs, /*deps=*/false, /*pack=*/false,
                             profile_, stream_, err);
  if (rc == RLIC_OK && !w64.fields.empty())
    rc = launch_queue<double>(w64.fields, w64.ksrc, w64.segs, /*deps=*/false, /*pack=*/false,
                              profile_, stream_, err);
  if (rc != RLIC_OK) return rc;  // side work still running is waited for by ~SideWork
  if (n_single > 0) {  // the step is synchronous
    hipError_t e = side.join();
    if (hipError_t se = hipStreamSynchronize(stream_); e == hipSuccess) e = se;
    if (e != hipSuccess) {
      (void)hipGetLastError();
      err = std::string("serving step: ") + hipGetErrorString(e);
      return RLIC_ERR_HIP;
    }
  }
  commit_dtype(w32, finished);
  commit_dtype(w64, finished);
  return RLIC_OK;
}

int Server::extend(int64_t handle, int64_t more, std::string& err) {
  std::lock_guard<std::mutex> lock(mu_);
  auto go = [&](auto& tab) -> int {
    auto it = tab.find(handle);
    if (it == tab.end()) return 1;
    auto& r = it->second;
    if (r.done < r.target) {
      err = "the field still has passes running";
      return RLIC_ERR_INVALID_ARGUMENT;
    }
    if (r.done > 0 && !r.recorded) {
      err = "the field keeps no move record: it cannot run more passes";
      return RLIC_ERR_INVALID_ARGUMENT;
    }
    r.target += more;
    return RLIC_OK;
  };
  int rc = go(f32_);
  if (rc == 1) rc = go(f64_);
  if (rc == 1) {
    err = "unknown field handle " + std::to_string(handle);
    return RLIC_ERR_INVALID_ARGUMENT;
  }
  return rc;
}

template <typename T>
int Server::read(int64_t handle, T* out, std::string& err) {
  std::lock_guard<std::mutex> lock(mu_);
  auto it = table<T>().find(handle);
  if (it == table<T>().end()) {
    err = "unknown field handle " + std::to_string(handle) + " for this dtype";
    return RLIC_ERR_INVALID_ARGUMENT;
  }
  const Resident<T>& r = it->second;
  const size_t n = static_cast<size_t>(r.g.ny) * r.g.nx;
  if (n == 0) return RLIC_OK;  // an empty field has an empty output
  if (!out) {
    err = "null output pointer";
    return RLIC_ERR_INVALID_ARGUMENT;
  }
  if (r.cur < 0) {
    err = "the field has no output yet";
    return RLIC_ERR_INVALID_ARGUMENT;
  }
  hipError_t e = hipMemcpyAsync(out, r.buf[r.cur], sizeof(T) * n, hipMemcpyDefault, stream_);
  if (e == hipSuccess) e = hipStreamSynchronize(stream_);
  if (e != hipSuccess) {
    err = std::string("reading field output: ") + hipGetErrorString(e);
    return RLIC_ERR_HIP;
  }
  return RLIC_OK;
}

int Server::release(int64_t handle, std::string& err) {
  std::lock_guard<std::mutex> lock(mu_);
  auto go = [&](auto& tab) -> bool {
    auto it = tab.find(handle);
    if (it == tab.end()) return false;
    used_ -= static_cast<int64_t>(it->second.bytes);
    give_block(it->second.block, it->second.block_bytes);
    tab.erase(it);
    return true;
  };
  if (go(f32_) || go(f64_)) return RLIC_OK;
  err = "unknown field handle " + std::to_string(handle);
  return RLIC_ERR_INVALID_ARGUMENT;
}

void Server::usage(int64_t* used_bytes, int64_t* resident, int64_t* active) {
  std::lock_guard<std::mutex> lock(mu_);
  int64_t act = 0;
  for (auto& kv : f32_) act += kv.second.done < kv.second.target;
  for (auto& kv : f64_) act += kv.second.done < kv.second.target;
  if (used_bytes) *used_bytes = used_;
  if (resident) *resident = static_cast<int64_t>(f32_.size() + f64_.size());
  if (active) *active = act;
}

#define RLIC_SERVER_INSTANTIATE(T)                                                               \
  template size_t Server::field_bytes<T>(int64_t, int64_t, int64_t, bool);                       \
  template int Server::admit<T>(const T*, int64_t, int64_t, const T*, int64_t, int64_t, const T*, \
                                int64_t, int64_t, int64_t, int64_t, const T*, int64_t, int, int, \
                                int, int, int, int64_t, bool, int64_t*, std::string&);           \
  template int Server::read<T>(int64_t, T*, std::string&);

RLIC_SERVER_INSTANTIATE(float)
RLIC_SERVER_INSTANTIATE(double)

}  // namespace rlic
