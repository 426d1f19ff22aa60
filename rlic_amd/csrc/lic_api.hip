// C ABI of librlic_hip.so (declared in include/rlic_hip.h).
//
// Iteration driver (src/lib.rs:408-443, convolve_iteratively): the
// reference copies the texture into a C-order buffer, runs `iterations`
// passes and swaps input/output between passes.  Here:
//   1. lic_pack builds the guarded (u, v, texture) cell array, a ring of
//      boundary-rule guard cells around the image, from the caller's
//      strided arrays (any strides, including transposed views);
//   2. lic_trace runs pass 1 and, when more passes follow, records the move
//      of every streamline step;
//   3. lic_replay runs passes 2..n from the recorded moves, ping-ponging
//      two device buffers so that the last pass lands in `out`.
// When the move record would not fit the memory budget, every pass is
// traced again instead (texture slot re-packed between passes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rlic_hip.h"
#include "lic_batch.h"
#include "lic_equalize.h"
#include "lic_launch.h"
#include "lic_scratch.h"
#include "lic_serve.h"
#include "lic_streams.h"

namespace {

thread_local std::string t_last_error;

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  t_last_error = buf;
  return code;
}

int hip_error(hipError_t e, const char* what) {
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
    return set_error(RLIC_ERR_OUT_OF_MEMORY, "%s: out of device memory (%s)", what,
                     hipGetErrorString(e));
  }
  return set_error(RLIC_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define RLIC_HIP_TRY(expr, what)                  \
  do {                                            \
    hipError_t e_ = (expr);                       \
    if (e_ != hipSuccess) return hip_error(e_, what); \
  } while (0)

// Optional kernel timing: when enabled, every kernel launch of the library
// is bracketed by two hipEvents recorded on the launch's own stream; the
// pairs are resolved and aggregated per kernel name on read-out.
struct TimedLaunch {
  const char* name;
  hipEvent_t start, stop;
  int64_t units;  // launches the interval stands for (a graph of several passes)
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<TimedLaunch> g_prof_pending;
std::map<std::string, std::pair<int64_t, double>> g_prof_done;  // name -> (count, total ms)

// Lane work of tracing passes (RLIC_TRACE_STATS=1): live and executed
// lane-steps summed over the calls since the last read.
std::mutex g_stats_mu;
int64_t g_stats_useful = 0, g_stats_executed = 0;

// Set while this thread runs an allocation-free call: no events are
// created for it (hipEventCreate may allocate).
thread_local bool t_noalloc = false;

struct LaunchTimer {
  const char* name;
  hipStream_t stream;
  hipEvent_t start = nullptr, stop = nullptr;
  bool on = false;
  int64_t units = 1;
  LaunchTimer(const char* n, hipStream_t s, int64_t u = 1) : name(n), stream(s), units(u) {
    if (t_noalloc) return;
    {
      std::lock_guard<std::mutex> lock(g_prof_mu);
      on = g_prof_on;
    }
    if (!on) return;
    if (hipEventCreate(&start) != hipSuccess || hipEventCreate(&stop) != hipSuccess ||
        hipEventRecord(start, stream) != hipSuccess) {
      on = false;
    }
  }
  ~LaunchTimer() {
    if (!on) return;
    if (hipEventRecord(stop, stream) != hipSuccess) return;
    std::lock_guard<std::mutex> lock(g_prof_mu);
    g_prof_pending.push_back({name, start, stop, units});
  }
};

// Stream-ordered device allocations from the device's default memory pool,
// released on the same stream when the scope ends.  The pool keeps freed
// memory cached for the next call (release threshold raised once per
// device).
void configure_pool(int dev) {
  static std::mutex mu;
  static bool done[256] = {};
  std::lock_guard<std::mutex> lock(mu);
  if (dev < 0 || dev >= 256 || done[dev]) return;
  hipMemPool_t pool;
  if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
    uint64_t threshold = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold);
  }
  done[dev] = true;
}

struct DevMem {
  void* ptr = nullptr;
  hipStream_t stream = nullptr;
  DevMem() = default;
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
  ~DevMem() {
    if (ptr) (void)hipFreeAsync(ptr, stream);
  }
  hipError_t alloc(size_t bytes, hipStream_t s) {
    stream = s;
    if (bytes == 0) bytes = 1;
    return hipMallocAsync(&ptr, bytes, s);
  }
};

// Element-offset extent of a strided 2-D view: [lo, hi].
void view_span(int64_t ny, int64_t nx, int64_t sy, int64_t sx, int64_t& lo, int64_t& hi) {
  lo = 0;
  hi = 0;
  const int64_t dy = (ny - 1) * sy, dx = (nx - 1) * sx;
  if (dy < 0) lo += dy; else hi += dy;
  if (dx < 0) lo += dx; else hi += dx;
}

int env_int(const char* name, int dflt) {
  const char* s = std::getenv(name);
  if (!s || !*s) return dflt;
  return std::atoi(s);
}

// How a call's passes 2..n replay: lds = 1, LDS-staged with a th x tw tile;
// lds = 2, the hybrid (64 x 64 tiles, the first `band` steps of each
// direction from an LDS box of that halo, the rest from the image); lds =
// 0, L2 gathers over a tile map (strip height `band`, wave shape).
struct ReplayPlan {
  int32_t lds, th, tw, band, wave_w_log2;
};

template <typename T>
hipError_t launch_plan(const ReplayPlan& p, const T* src, T* dst, const T* kernel, int32_t klen,
                       const rlic::FieldGeom& g, int profile, const uint4* codes,
                       hipStream_t stream) {
  if (p.lds == 2)
    return rlic::launch_replay_lds_hybrid<T>(src, kernel, klen, g, p.band, profile, dst, codes,
                                             stream);
  if (p.lds)
    return rlic::launch_replay_lds_tile<T>(src, kernel, klen, g, rlic::LdsTile{p.th, p.tw},
                                           profile, dst, codes, stream);
  const rlic::TileMap m = rlic::make_tile_map(g.ny, g.nx, p.band, p.wave_w_log2);
  return rlic::launch_replay<T>(src, kernel, klen, g, m, profile, dst, codes, stream);
}

const char* plan_kernel_name(const ReplayPlan& p) { return p.lds ? "lic_replay_lds" : "lic_replay"; }

// LDS plans stage a box per tile (192 x 192 for a 64 x 64 float tile at
// kmid 64) whatever the image size.  Below 2^18 pixels (512 x 512) the
// staging outweighs the gathers it saves: per-call loops over fields of
// 2^12..2^18 pixels ran 1.1-1.5x faster with L2 gathers (8.7 vs 13.4 ms,
// 12.9 vs 17.1 ms, 13.9 vs 15.2 ms by size class), and the hybrid doubled
// a small-field loop, so smaller images use the L2-gather replay.
// RLIC_REPLAY_LDS (whole boxes and hybrid) and RLIC_REPLAY_HYBRID: 0 never,
// 1 (default) from 2^18 pixels, 2 at any size (tests).
constexpr int64_t kLdsMinPixels = int64_t{1} << 18;

bool lds_allowed(int64_t n) {
  const int mode = env_int("RLIC_REPLAY_LDS", 1);
  return mode != 0 && (n >= kLdsMinPixels || mode == 2);
}

template <typename T>
bool hybrid_allowed(int32_t klen, int64_t n, int32_t halo) {
  const int mode = env_int("RLIC_REPLAY_HYBRID", 1);
  return mode != 0 && (lds_allowed(n) || (mode == 2 && env_int("RLIC_REPLAY_LDS", 1) != 0)) &&
         rlic::lds_hybrid_fits<T>(klen, n, halo);
}

// The hybrid's halo: the largest whose box fits (64 in float, 32 in
// double), when kmid exceeds it; 0 otherwise.  A 32-cell halo in float lost
// to the whole 64x64 box at every kernel length measured (L = 67..129 at
// 4096^2: 0.48 vs 0.40 ms per pass at L=67, 1.08 vs 0.48 at L=129) and to
// the 64-cell halo beyond (L=199: 2.28 vs 1.51 ms).
template <typename T>
int32_t hybrid_halo(int32_t klen, int64_t n) {
  const int32_t halo = rlic::hybrid_max_halo<T>();
  return hybrid_allowed<T>(klen, n, halo) ? halo : 0;
}

// Default plan: the hybrid when it applies (kmid above its halo, an image
// of at least 2^18 pixels), else the largest LDS tile whose whole box fits
// (from 2^18 pixels), else L2 gathers over the tracing pass's tile map.  Measured per pass (4096^2 float,
// 2048^2 double): the hybrid beat every whole box, e.g. float L=135 0.59
// vs 0.64 ms (64x64 box), L=151 0.76 vs 1.27 ms (32x32), double L=75 0.12
// vs 0.15 ms (64x64): its box is smaller and only the last few steps of a
// line leave it.
template <typename T>
ReplayPlan default_plan(int32_t klen, int64_t n) {
  if (const int32_t halo = hybrid_halo<T>(klen, n)) return {2, 64, 64, halo, 0};
  const rlic::LdsTile t = rlic::lds_replay_tile<T>(klen);
  if (lds_allowed(n) && t.h > 0) return {1, t.h, t.w, 0, 0};
  return {0, 0, 0, env_int("RLIC_TILE_BAND", 16), env_int("RLIC_WAVE_W_LOG2", 3)};
}

// Replay autotuner.  The first call with at least 3 passes on a field of at
// least 2^18 pixels for a new (dtype, shape, kernel length, profile) times
// its pass 2 under every candidate plan — each LDS tile whose box fits
// (64x64, 32x64, 32x32), the hybrid when it applies, and L2 gathers with
// two strip heights — on the
// caller's stream with HIP events (all candidates write the same bits), and
// keeps the fastest for that key.  RLIC_AUTOTUNE=0 disables it;
// RLIC_AUTOTUNE_CACHE=<file> loads and appends the table as text.
struct TuneKey {
  int32_t dtype_size, ny, nx, klen, profile;
  bool operator<(const TuneKey& o) const { return std::memcmp(this, &o, sizeof(*this)) < 0; }
};
struct TuneEntry {
  ReplayPlan plan;
  float ms;
};
struct Tuner {
  std::mutex mu;
  std::map<TuneKey, TuneEntry> table;
  bool loaded = false;
};
Tuner g_tuner;

// Adds the plan lines of a text file to the table; the number of lines read,
// or -1 when the file cannot be opened.
int64_t tuner_read_file_locked(const char* path) {
  FILE* f = std::fopen(path, "r");
  if (!f) return -1;
  TuneKey k;
  TuneEntry e;
  int64_t count = 0;
  while (std::fscanf(f, " rlic-replay-plan %d %d %d %d %d %d %d %d %d %d %f", &k.dtype_size,
                     &k.ny, &k.nx, &k.klen, &k.profile, &e.plan.lds, &e.plan.th, &e.plan.tw,
                     &e.plan.band, &e.plan.wave_w_log2, &e.ms) == 11) {
    g_tuner.table[k] = e;
    ++count;
  }
  std::fclose(f);
  return count;
}

void tuner_load_locked() {
  if (g_tuner.loaded) return;
  g_tuner.loaded = true;
  const char* path = std::getenv("RLIC_AUTOTUNE_CACHE");
  if (path && *path) (void)tuner_read_file_locked(path);
}

// A plan from the table is used only if it can run here.
template <typename T>
bool plan_usable(const ReplayPlan& p, int32_t klen, int64_t n) {
  if (p.lds == 2) return p.th == 64 && p.tw == 64 && rlic::lds_hybrid_fits<T>(klen, n, p.band);
  if (p.lds) {
    const rlic::LdsTile big = rlic::lds_replay_tile<T>(klen);
    const bool shape = (p.th == 64 && p.tw == 64) || (p.th == 32 && (p.tw == 64 || p.tw == 32));
    if (!shape || big.h == 0) return false;
    return p.th * p.tw <= big.h * big.w;  // boxes of smaller tiles are smaller
  }
  return p.band >= 1 && p.band <= 1024 && p.wave_w_log2 >= 3 && p.wave_w_log2 <= 5;
}

void tuner_store_locked(const TuneKey& k, const TuneEntry& e) {
  g_tuner.table[k] = e;
  const char* path = std::getenv("RLIC_AUTOTUNE_CACHE");
  if (!path || !*path) return;
  if (FILE* f = std::fopen(path, "a")) {
    std::fprintf(f, "rlic-replay-plan %d %d %d %d %d %d %d %d %d %d %.6f\n", k.dtype_size, k.ny,
                 k.nx, k.klen, k.profile, e.plan.lds, e.plan.th, e.plan.tw, e.plan.band,
                 e.plan.wave_w_log2, e.ms);
    std::fclose(f);
  }
}

template <typename T>
std::vector<ReplayPlan> replay_candidates(int32_t klen, int64_t n) {
  std::vector<ReplayPlan> c;
  const rlic::LdsTile big = rlic::lds_replay_tile<T>(klen);
  if (env_int("RLIC_REPLAY_LDS", 1) != 0 && big.h > 0) {
    if (big.h == 64) c.push_back({1, 64, 64, 0, 0});
    if (big.h == 64 || big.w == 64) c.push_back({1, 32, 64, 0, 0});
    c.push_back({1, 32, 32, 0, 0});
  }
  if (const int32_t halo = hybrid_halo<T>(klen, n)) c.push_back({2, 64, 64, halo, 0});
  c.push_back({0, 0, 0, 16, 3});
  c.push_back({0, 0, 0, 4, 3});
  return c;
}

// Pass 2 under every candidate (src -> dst), timed; returns the fastest.
template <typename T>
hipError_t tune_replay(const std::vector<ReplayPlan>& cands, const T* src, T* dst,
                       const T* kernel, int32_t klen, const rlic::FieldGeom& g, int profile,
                       const uint4* codes, hipStream_t stream, TuneEntry& best) {
  std::vector<hipEvent_t> ev(cands.size() + 1, nullptr);
  hipError_t e = hipSuccess;
  for (auto& x : ev)
    if (e == hipSuccess) e = hipEventCreate(&x);
  if (e == hipSuccess) e = hipEventRecord(ev[0], stream);
  for (size_t k = 0; k < cands.size() && e == hipSuccess; ++k) {
    {
      LaunchTimer timer(plan_kernel_name(cands[k]), stream);
      e = launch_plan<T>(cands[k], src, dst, kernel, klen, g, profile, codes, stream);
    }
    if (e == hipSuccess) e = hipEventRecord(ev[k + 1], stream);
  }
  if (e == hipSuccess) e = hipEventSynchronize(ev.back());
  best.ms = -1.f;
  for (size_t k = 0; k < cands.size() && e == hipSuccess; ++k) {
    float ms = 0.f;
    e = hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
    if (e == hipSuccess && (best.ms < 0.f || ms < best.ms)) best = {cands[k], ms};
  }
  for (auto& x : ev)
    if (x) (void)hipEventDestroy(x);
  return e;
}

// The replay passes of one call as one hipGraph.  The passes are captured
// on a private stream (capture records, it runs nothing), instantiated, and
// launched on the caller's stream; instantiated graphs are cached by every
// argument of their launches (buffers included), so a call that repeats an
// earlier one on the same buffers replays the instantiated graph.
struct ReplayKey {
  int32_t dtype_size, device;
  const void* a;
  const void* b;
  const void* kernel;
  const void* codes;
  int32_t klen;
  rlic::FieldGeom g;
  ReplayPlan plan;
  int32_t profile, pad;  // pad: no implicit padding bytes (compared with memcmp)
  int64_t iterations;
  bool operator<(const ReplayKey& o) const { return std::memcmp(this, &o, sizeof(*this)) < 0; }
};

struct GraphCache {
  std::mutex mu;
  std::map<ReplayKey, std::pair<hipGraphExec_t, uint64_t>> entries;  // exec, last use
  uint64_t clock = 0;
  hipStream_t capture[256] = {};
  static constexpr size_t kMax = 16;
};
GraphCache g_graphs;

template <typename T>
hipError_t launch_replays(const ReplayPlan& plan, const T* kernel, int32_t klen,
                          const rlic::FieldGeom& g, int profile, T* a, T* b, const uint4* codes,
                          int64_t first, int64_t iterations, hipStream_t stream) {
  // pass `it` (1-based) reads A when even, B when odd, and writes the other
  for (int64_t it = first; it <= iterations; ++it) {
    const T* src = (it % 2 == 0) ? a : b;
    T* dst = (it % 2 == 0) ? b : a;
    const hipError_t e = launch_plan<T>(plan, src, dst, kernel, klen, g, profile, codes, stream);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Runs the replay passes as a (cached) graph; false when graphs are not
// available here (the caller then launches the passes directly).
template <typename T>
bool replay_graph(int device, const ReplayPlan& plan, const T* kernel, int32_t klen,
                  const rlic::FieldGeom& g, int profile, T* a, T* b, const uint4* codes,
                  int64_t iterations, hipStream_t stream) {
  if (device < 0 || device >= 256) return false;
  ReplayKey key;
  std::memset(&key, 0, sizeof(key));
  key.dtype_size = sizeof(T);
  key.device = device;
  key.a = a;
  key.b = b;
  key.kernel = kernel;
  key.codes = codes;
  key.klen = klen;
  key.g = g;
  key.plan = plan;
  key.profile = profile;
  key.iterations = iterations;
  std::lock_guard<std::mutex> lock(g_graphs.mu);
  hipGraphExec_t exec = nullptr;
  auto it = g_graphs.entries.find(key);
  if (it != g_graphs.entries.end()) {
    exec = it->second.first;
    it->second.second = ++g_graphs.clock;
  } else {
    hipStream_t& cap = g_graphs.capture[device];
    if (!cap && hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess) {
      (void)hipGetLastError();
      cap = nullptr;
      return false;
    }
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    const hipError_t le = launch_replays<T>(plan, kernel, klen, g, profile, a, b, codes, 2,
                                            iterations, cap);
    const hipError_t ce = hipStreamEndCapture(cap, &graph);
    if (le != hipSuccess || ce != hipSuccess || !graph ||
        hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGetLastError();
      if (graph) (void)hipGraphDestroy(graph);
      return false;
    }
    (void)hipGraphDestroy(graph);
    {
      std::lock_guard<std::mutex> plock(g_prof_mu);
      if (g_prof_on) g_prof_done["lic_graph_capture"].first += 1;  // captures, as a count
    }
    if (g_graphs.entries.size() >= GraphCache::kMax) {  // drop the least recently used
      auto lru = g_graphs.entries.begin();
      for (auto e = g_graphs.entries.begin(); e != g_graphs.entries.end(); ++e)
        if (e->second.second < lru->second.second) lru = e;
      (void)hipGraphExecDestroy(lru->second.first);
      g_graphs.entries.erase(lru);
    }
    g_graphs.entries.emplace(key, std::make_pair(exec, ++g_graphs.clock));
  }
  LaunchTimer timer(plan_kernel_name(plan), stream, iterations - 1);
  if (hipGraphLaunch(exec, stream) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return true;
}

// One call's scratch block: padded weights | cells | ping-pong buffer | move
// record (the cells a guarded array: a ring of guard cells around the
// image).  `codes` is where the move record starts; `codes + code_bytes`
// is the whole block with the record, `codes` without it.
struct ScratchLayout {
  size_t cells, tmp, codes, code_bytes;
};

template <typename T>
ScratchLayout scratch_layout(int32_t ny, int32_t nx, int32_t klen, bool need_tmp) {
  auto up = [](size_t x) { return (x + 255) & ~size_t{255}; };
  const size_t n = static_cast<size_t>(ny) * static_cast<size_t>(nx);
  const size_t ncells = static_cast<size_t>(rlic::guarded_cells(ny, nx));
  ScratchLayout l;
  l.cells = up(sizeof(T) * static_cast<size_t>(klen + 2 * rlic::kernel_pad()));
  l.tmp = l.cells + up(sizeof(rlic::Cell<T>) * ncells);
  l.codes = l.tmp + (need_tmp ? up(sizeof(T) * n) : 0);
  // code_chunks(klen) 16-byte chunks per pixel
  l.code_bytes = sizeof(uint4) * static_cast<size_t>(rlic::code_chunks(klen)) * n;
  return l;
}

// A caller-provided scratch region (the allocation-free entry points):
// `bytes` at `ptr`, device memory of the current device.
struct FixedScratch {
  void* ptr;
  size_t bytes;
};

// The paired trace (two lanes per pixel, one per direction) where a field
// has too few pixels to fill the GPU with lic_trace's one lane per pixel:
// up to RLIC_TRACE_PAIR_MAX pixels (0: never), when it applies.
constexpr int kTracePairMaxPixels = 1 << 17;

template <typename T>
bool use_trace_pair(int32_t klen, int32_t ny, int32_t nx) {
  const int64_t max_px = env_int("RLIC_TRACE_PAIR_MAX", kTracePairMaxPixels);
  return static_cast<int64_t>(ny) * nx <= max_px && rlic::trace_pair_fits<T>(klen, ny, nx);
}

// `fixed`: take the call's scratch from that region and allocate nothing
// (no move record when it does not fit: every pass is traced; no trace
// statistics, autotuning, graph capture or launch timing, which create HIP
// objects); nullptr: from the per-stream cache or the stream-ordered pool.
template <typename T>
int run_on_device(const T* tex, int64_t tsy, int64_t tsx, const T* u, int64_t usy, int64_t usx,
                  const T* v, int64_t vsy, int64_t vsx, int32_t ny, int32_t nx,
                  const T* kernel, int32_t klen, int uv_mode, const rlic::FieldGeom& g,
                  int64_t iterations, T* out, int profile, hipStream_t stream,
                  const FixedScratch* fixed = nullptr) {
  if (iterations == 0) {
    LaunchTimer timer("lic_copy2d", stream);
    RLIC_HIP_TRY(rlic::launch_copy2d<T>(tex, tsy, tsx, ny, nx, out, stream), "copy");
    return RLIC_OK;
  }
  const int64_t n = static_cast<int64_t>(ny) * nx;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    dev = -1;
  }
  const int32_t pad = rlic::kernel_pad();
  const bool stats = !fixed && env_int("RLIC_TRACE_STATS", 0) != 0;
  const bool need_tmp = iterations > 1 || stats;
  const ScratchLayout lay = scratch_layout<T>(ny, nx, klen, need_tmp);
  const size_t code_bytes = lay.code_bytes;
  bool replay = need_tmp && env_int("RLIC_REPLAY", 1) != 0;
  if (!fixed && replay && code_bytes > (size_t{256} << 20)) {
    // a large record must leave half of the free memory; small ones skip
    // the (slow) query and fall back below if the allocation fails
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
    replay = code_bytes <= free_b / 2;
  }
  const size_t off_cells = lay.cells, off_tmp = lay.tmp, off_codes = lay.codes;
  rlic::CallScratch scratch;
  char* base;
  if (fixed) {
    if (off_codes > fixed->bytes)
      return set_error(RLIC_ERR_INVALID_ARGUMENT,
                       "the call needs %zu bytes of scratch; the fixed scratch holds %zu",
                       off_codes, fixed->bytes);
    if (replay && off_codes + code_bytes > fixed->bytes) replay = false;  // trace every pass
    base = static_cast<char*>(fixed->ptr);
  } else {
    hipError_t ae = scratch.get(off_codes + (replay ? code_bytes : 0), stream);
    if (ae != hipSuccess && replay) {  // no room for the record: trace every pass
      (void)hipGetLastError();
      replay = false;
      ae = scratch.get(off_codes, stream);
    }
    RLIC_HIP_TRY(ae, "allocating call scratch");
    base = static_cast<char*>(scratch.ptr());
  }
  T* kpad = reinterpret_cast<T*>(base);
  auto* c = reinterpret_cast<rlic::Cell<T>*>(base + off_cells);
  T* tmp = need_tmp ? reinterpret_cast<T*>(base + off_tmp) : nullptr;
  uint4* code_ptr = replay ? reinterpret_cast<uint4*>(base + off_codes) : nullptr;
  {
    // the cells and the padded weight copy in one launch
    LaunchTimer timer("lic_pack", stream);
    RLIC_HIP_TRY(rlic::launch_pack<T>(tex, tsy, tsx, u, usy, usx, v, vsy, vsx, g, c, stream,
                                      kernel, klen, kpad),
                 "pack");
  }
  kernel = kpad + pad;
  // trace tiles of 32 x 8 pixels (16 x 4 per wave) in XCD bands of 64
  // tiles: 2.57 vs 2.65 ms per trace launch at 4096^2, L=129 (sweep of
  // bands 4..64, wave widths 8..32 pixels)
  const rlic::TileMap m = rlic::make_tile_map(ny, nx, env_int("RLIC_TILE_BAND", 64),
                                              env_int("RLIC_WAVE_W_LOG2", 4));
  // one tracing pass into dst, recording the moves into codes (or not)
  const bool pair = !stats && use_trace_pair<T>(klen, ny, nx);
  auto trace = [&](T* dst, uint4* codes) {
    LaunchTimer timer(pair ? "lic_trace_pair" : "lic_trace", stream);
    return pair ? rlic::launch_trace_pair<T>(c, kernel, klen, g, m, uv_mode, profile, dst, codes,
                                             stream)
                : rlic::launch_trace<T>(c, kernel, klen, g, m, uv_mode, profile, dst, codes,
                                        stream);
  };
  if (iterations == 1 && !stats) {
    RLIC_HIP_TRY(trace(out, nullptr), "trace");
    return RLIC_OK;
  }
  if (replay) {
    uint4* codes = code_ptr;
    // pass `it` (1-based) writes A when odd, B when even; the last lands in out
    T* a = ((iterations - 1) % 2 == 0) ? out : tmp;
    T* b = (a == out) ? tmp : out;
    RLIC_HIP_TRY(trace(a, codes), "trace");
    if (stats) {  // measurement: lane work of the pass, from its move record
      DevMem st;
      unsigned long long host_st[2] = {0, 0};
      RLIC_HIP_TRY(st.alloc(sizeof(host_st), stream), "allocating trace statistics");
      RLIC_HIP_TRY(hipMemsetAsync(st.ptr, 0, sizeof(host_st), stream), "clearing statistics");
      RLIC_HIP_TRY(rlic::launch_record_stats(codes, klen, g, m,
                                             static_cast<unsigned long long*>(st.ptr), stream),
                   "trace statistics");
      RLIC_HIP_TRY(hipMemcpyAsync(host_st, st.ptr, sizeof(host_st), hipMemcpyDeviceToHost, stream),
                   "reading statistics");
      RLIC_HIP_TRY(hipStreamSynchronize(stream), "reading statistics");
      std::lock_guard<std::mutex> lock(g_stats_mu);
      g_stats_useful += static_cast<int64_t>(host_st[0]);
      g_stats_executed += static_cast<int64_t>(host_st[1]);
    }
    if (iterations == 1) return RLIC_OK;  // recorded only for the statistics
    // the replay plan: tuned (see Tuner), or the default
    ReplayPlan plan = default_plan<T>(klen, n);
    int64_t first = 2;
    if (env_int("RLIC_AUTOTUNE", 1) != 0 && iterations >= 3 && n >= (int64_t{1} << 18)) {
      // an allocation-free call uses a tuned plan when the table has one,
      // and never tunes (the timing events are HIP objects)
      const TuneKey key{static_cast<int32_t>(sizeof(T)), ny, nx, klen, profile};
      std::lock_guard<std::mutex> lock(g_tuner.mu);
      tuner_load_locked();
      auto hit = g_tuner.table.find(key);
      if (hit != g_tuner.table.end()) {
        if (plan_usable<T>(hit->second.plan, klen, n)) plan = hit->second.plan;
      } else if (!fixed) {
        const std::vector<ReplayPlan> cands = replay_candidates<T>(klen, n);
        TuneEntry best{plan, -1.f};
        RLIC_HIP_TRY(tune_replay<T>(cands, a, b, kernel, klen, g, profile, codes, stream, best),
                     "tuning the replay");
        tuner_store_locked(key, best);
        plan = best.plan;
        first = 3;  // pass 2 is done (every candidate wrote the same bits)
      }
    }
    // passes first..n: direct launches, or one captured graph with
    // RLIC_GRAPH=1 (measured no faster on MI355X: kept as an option)
    if (first == 2 && !fixed && env_int("RLIC_GRAPH", 0) != 0 &&
        replay_graph<T>(dev, plan, kernel, klen, g, profile, a, b, codes, iterations, stream))
      return RLIC_OK;
    for (int64_t it = first; it <= iterations; ++it) {
      const T* src = (it % 2 == 0) ? a : b;
      T* dst = (it % 2 == 0) ? b : a;
      LaunchTimer timer(plan_kernel_name(plan), stream);
      RLIC_HIP_TRY(launch_plan<T>(plan, src, dst, kernel, klen, g, profile, codes, stream),
                   "replay");
    }
    return RLIC_OK;
  }
  for (int64_t it = 1; it <= iterations; ++it) {
    T* dst = (it == iterations) ? out : tmp;
    RLIC_HIP_TRY(trace(dst, nullptr), "trace");
    if (it != iterations) {
      LaunchTimer timer("lic_repack_texture", stream);
      RLIC_HIP_TRY(rlic::launch_repack_texture<T>(tmp, g, c, stream),
                   "repack");
    }
  }
  return RLIC_OK;
}

// Size limits of a non-empty image (ny, nx >= 1), checked before any
// product is formed: each side and the pixel count below 2^31 (positions
// and pixel indices are 32-bit), and the guarded cell array of
// (ny + 2) x (nx + 2) cells addressable by a 32-bit cell index (thin images,
// e.g. 1 x N with N > 1,431,655,763, pass the pixel limit but not this one).
int check_size(int64_t ny, int64_t nx, int64_t klen) {
  if (ny > INT32_MAX || nx > INT32_MAX || ny * nx > INT32_MAX)
    return set_error(RLIC_ERR_INVALID_ARGUMENT,
                     "image of %lld x %lld pixels exceeds the supported 2^31-1 pixels",
                     static_cast<long long>(ny), static_cast<long long>(nx));
  if (rlic::guarded_cells(ny, nx) > int64_t{UINT32_MAX})
    return set_error(RLIC_ERR_INVALID_ARGUMENT,
                     "image of %lld x %lld pixels exceeds the supported size: its guarded "
                     "array of %lld x %lld cells exceeds 2^32-1 cells",
                     static_cast<long long>(ny), static_cast<long long>(nx),
                     static_cast<long long>(ny + 2), static_cast<long long>(nx + 2));
  if (klen > INT32_MAX)
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "kernel of %lld weights exceeds 2^31-1",
                     static_cast<long long>(klen));
  return RLIC_OK;
}

// Argument checks shared by the single and batched entry points.
template <typename T>
int validate(const T* tex, const T* u, const T* v, int64_t ny, int64_t nx, const T* kernel,
             int64_t klen, int uv_mode, int xl, int xr, int yl, int yr, int64_t iterations,
             const T* out) {
  if (ny < 0 || nx < 0) return set_error(RLIC_ERR_INVALID_ARGUMENT, "negative image size");
  if (klen < 0) return set_error(RLIC_ERR_INVALID_ARGUMENT, "negative kernel length");
  if (iterations < 0)
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "Invalid number of iterations: %lld",
                     static_cast<long long>(iterations));
  if (uv_mode != RLIC_UV_VELOCITY && uv_mode != RLIC_UV_POLARIZATION)
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "unknown uv_mode %d", uv_mode);
  for (int b : {xl, xr, yl, yr}) {
    if (b != RLIC_BOUND_CLOSED && b != RLIC_BOUND_PERIODIC)
      return set_error(RLIC_ERR_INVALID_ARGUMENT, "unknown boundary %d", b);
  }
  if (ny == 0 || nx == 0) return RLIC_OK;  // nothing to compute (the reference loops zero times)
  if (const int rc = check_size(ny, nx, klen); rc != RLIC_OK) return rc;
  if (!tex || !u || !v || !out) return set_error(RLIC_ERR_INVALID_ARGUMENT, "null array pointer");
  if (iterations > 0 && klen == 0)
    return set_error(RLIC_ERR_EMPTY_KERNEL,
                     "kernel is empty: kernel[len(kernel) // 2] does not exist");
  if (klen > 0 && !kernel) return set_error(RLIC_ERR_INVALID_ARGUMENT, "null kernel pointer");
  return RLIC_OK;
}

// Host-mapped pinned words for the value checks' results, one buffer per
// host thread, grown on demand and held until the process ends (a call
// synchronizes before returning, so the next call of the thread may reuse
// them).  The kernels write the flags there directly: no device-to-host
// copy, no memset.
struct PinnedWords {
  char* host = nullptr;
  char* dev = nullptr;  // the same memory's device address (portable: every device)
  size_t bytes = 0;
};
thread_local PinnedWords t_pinned;

hipError_t pinned_words(size_t bytes, char** host, char** dev) {
  if (t_pinned.bytes < bytes) {
    size_t want = 4096;
    while (want < bytes) want *= 2;
    if (t_pinned.host) (void)hipHostFree(t_pinned.host);
    t_pinned = PinnedWords{};
    void* h = nullptr;
    if (hipError_t e = hipHostMalloc(&h, want, hipHostMallocMapped | hipHostMallocPortable);
        e != hipSuccess)
      return e;
    void* d = nullptr;
    if (hipError_t e = hipHostGetDevicePointer(&d, h, 0); e != hipSuccess) {
      (void)hipHostFree(h);
      return e;
    }
    t_pinned = PinnedWords{static_cast<char*>(h), static_cast<char*>(d), want};
  }
  *host = t_pinned.host;
  *dev = t_pinned.dev;
  return hipSuccess;
}

// The element checks of rlic.convolve (src/rlic/_lib.py:174, :194) for
// `n` fields (sizes already validated) in one launch: *first_bad = the
// first field whose texture holds an element below 0 or whose kernel a
// non-finite weight, else -1; *what (if not null) = its problems (bit 0:
// texture, bit 1: kernel).
template <typename T>
int check_values_run(const rlic::CheckField<T>* fields, int64_t n, int64_t* first_bad, int* what,
                     void* stream_ptr) {
  *first_bad = -1;
  if (what) *what = 0;
  int64_t max_pixels = 0;
  bool any_values = false;
  for (int64_t k = 0; k < n; ++k) {
    const rlic::CheckField<T>& f = fields[k];
    const int64_t pixels = f.ny > 0 && f.nx > INT64_MAX / f.ny ? INT64_MAX : f.ny * f.nx;
    max_pixels = std::max(max_pixels, pixels);
    any_values = any_values || pixels > 0 || f.kernel_len > 0;
  }
  if (!any_values) return RLIC_OK;  // nothing to read: no device work
  hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
  const bool inline_fields = n <= rlic::kCheckInline;
  const size_t flag_bytes = (sizeof(uint32_t) * 2 * static_cast<size_t>(n) + 255) & ~size_t{255};
  const size_t table_bytes = inline_fields ? 0 : sizeof(rlic::CheckField<T>) * static_cast<size_t>(n);
  char* host = nullptr;
  char* dev = nullptr;
  RLIC_HIP_TRY(pinned_words(flag_bytes + table_bytes, &host, &dev), "allocating the check flags");
  auto* flags = reinterpret_cast<uint32_t*>(host);
  std::memset(flags, 0, sizeof(uint32_t) * 2 * static_cast<size_t>(n));
  auto* dev_flags = reinterpret_cast<uint32_t*>(dev);
  rlic::CallScratch scratch;
  if (inline_fields) {
    rlic::CheckPack<T> pack{};
    for (int64_t k = 0; k < n; ++k) pack.f[k] = fields[k];
    RLIC_HIP_TRY(rlic::launch_check_values<T>(nullptr, &pack, static_cast<int32_t>(n), max_pixels,
                                              dev_flags, stream),
                 "checking values");
  } else {
    std::memcpy(host + flag_bytes, fields, table_bytes);
    RLIC_HIP_TRY(scratch.get(table_bytes, stream), "allocating the check's field table");
    RLIC_HIP_TRY(hipMemcpyAsync(scratch.ptr(), host + flag_bytes, table_bytes,
                                hipMemcpyHostToDevice, stream),
                 "uploading the field table");
    RLIC_HIP_TRY(rlic::launch_check_values<T>(static_cast<const rlic::CheckField<T>*>(scratch.ptr()),
                                              nullptr, static_cast<int32_t>(n), max_pixels,
                                              dev_flags, stream),
                 "checking values");
  }
  RLIC_HIP_TRY(hipStreamSynchronize(stream), "checking values");
  const volatile uint32_t* done_flags = flags;  // written by the device since the memset
  for (int64_t k = 0; k < n; ++k) {
    const uint32_t bits = (done_flags[2 * k] ? 1u : 0u) | (done_flags[2 * k + 1] ? 2u : 0u);
    if (bits) {
      *first_bad = k;
      if (what) *what = static_cast<int>(bits);
      break;
    }
  }
  return RLIC_OK;
}

// A batch's fields (only texture, strides, sizes and kernel are read).
template <typename T, typename FieldSpec>
int batch_check_entry(const FieldSpec* fields, int64_t nfields, int64_t* first_bad, int* what,
                      void* stream_ptr) {
  t_last_error.clear();
  if (!first_bad) return set_error(RLIC_ERR_INVALID_ARGUMENT, "null first_bad pointer");
  *first_bad = -1;
  if (what) *what = 0;
  if (nfields < 0 || (nfields > 0 && !fields))
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "invalid field list");
  if (nfields > INT32_MAX) return set_error(RLIC_ERR_INVALID_ARGUMENT, "too many fields");
  std::vector<rlic::CheckField<T>> table(static_cast<size_t>(nfields));
  for (int64_t k = 0; k < nfields; ++k) {
    const FieldSpec& f = fields[k];
    if (f.ny < 0 || f.nx < 0 || f.kernel_len < 0)
      return set_error(RLIC_ERR_INVALID_ARGUMENT, "field %lld: negative size",
                       static_cast<long long>(k));
    const bool empty = f.ny == 0 || f.nx == 0;  // any size: 64-bit indexing throughout
    if (!empty && !f.texture)
      return set_error(RLIC_ERR_INVALID_ARGUMENT, "field %lld: null texture pointer",
                       static_cast<long long>(k));
    if (f.kernel_len > 0 && !f.kernel)
      return set_error(RLIC_ERR_INVALID_ARGUMENT, "field %lld: invalid kernel",
                       static_cast<long long>(k));
    table[static_cast<size_t>(k)] = rlic::CheckField<T>{f.texture, f.tex_sy, f.tex_sx,
                                                        empty ? 0 : f.ny, empty ? 0 : f.nx,
                                                        f.kernel, f.kernel_len};
  }
  return check_values_run<T>(table.data(), nfields, first_bad, what, stream_ptr);
}

// One field's element checks.
template <typename T>
int check_entry(const T* texture, int64_t ny, int64_t nx, int64_t sy, int64_t sx, const T* kernel,
                int64_t kernel_len, int* what, void* stream_ptr) {
  t_last_error.clear();
  if (!what) return set_error(RLIC_ERR_INVALID_ARGUMENT, "null result pointer");
  *what = 0;
  if (ny < 0 || nx < 0 || kernel_len < 0)
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "negative size");
  const bool empty = ny == 0 || nx == 0;  // any size: 64-bit indexing throughout
  if (!empty && !texture) return set_error(RLIC_ERR_INVALID_ARGUMENT, "null texture pointer");
  if (kernel_len > 0 && !kernel)
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "invalid kernel");
  const rlic::CheckField<T> f{texture, sy, sx, empty ? 0 : ny, empty ? 0 : nx, kernel, kernel_len};
  int64_t first_bad = -1;
  return check_values_run<T>(&f, 1, &first_bad, what, stream_ptr);
}

template <typename T, typename FieldSpec>
int batch_entry(const FieldSpec* fields, int64_t nfields, int profile, void* stream_ptr) {
  t_last_error.clear();
  if (nfields < 0 || (nfields > 0 && !fields))
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "invalid field list");
  if (profile < 0 || profile > 3)
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "unknown arithmetic profile %d", profile);
  if (nfields > INT32_MAX) return set_error(RLIC_ERR_INVALID_ARGUMENT, "too many fields");
  for (int64_t k = 0; k < nfields; ++k) {
    const FieldSpec& f = fields[k];
    const int rc = validate<T>(f.texture, f.u, f.v, f.ny, f.nx, f.kernel, f.kernel_len,
                               f.uv_mode, f.x_left, f.x_right, f.y_left, f.y_right,
                               f.iterations, f.out);
    if (rc != RLIC_OK) {
      t_last_error = "field " + std::to_string(k) + ": " + t_last_error;
      return rc;
    }
    if (f.iterations > INT32_MAX)
      return set_error(RLIC_ERR_INVALID_ARGUMENT, "field %lld: too many iterations",
                       static_cast<long long>(k));
  }
  if (nfields == 0) return RLIC_OK;
  int dev = 0;
  RLIC_HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
  configure_pool(dev);
  hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
  // Fields of at least RLIC_BATCH_SINGLE_MIN pixels (default 2^18; 0 or
  // less: none) run as single-field launches (trace, then LDS-staged or
  // tuned replays); the rest share one persistent launch.  Either way a
  // field gives the bits of its single call.
  const int single_min = env_int("RLIC_BATCH_SINGLE_MIN", 1 << 18);
  std::vector<FieldSpec> rest;
  rest.reserve(static_cast<size_t>(nfields));
  std::vector<int64_t> big;
  for (int64_t k = 0; k < nfields; ++k) {
    const FieldSpec& f = fields[k];
    // an empty image has an empty output: no work, and its other side may
    // be any size (no per-field state is sized from it)
    if (f.ny == 0 || f.nx == 0) continue;
    if (single_min <= 0 || f.ny * f.nx < single_min) rest.push_back(f);
    else big.push_back(k);
  }
  // Single-field launches of a field under about 2^20 pixels fill only part
  // of the GPU and leave it idle in their tails, so the big fields are
  // spread over up to RLIC_BATCH_STREAMS (default 4; 1: the caller's stream
  // only) side streams, largest first to the least loaded one, where they
  // overlap with each other and with the persistent launch on the caller's
  // stream.
  const int64_t want_side =
      std::min<int64_t>(std::max(1, std::min(env_int("RLIC_BATCH_STREAMS", 4), 16)),
                        static_cast<int64_t>(big.size()));
  rlic::SideWork side(stream);
  if (want_side > 1) (void)side.start(dev, static_cast<int>(want_side));
  rlic::BatchScratch scratch;
  std::string err;
  int rc = RLIC_OK;
  {
    LaunchTimer timer("lic_batch", stream);
    std::vector<int> lane_of(big.size(), 0);
    if (!side.streams.empty()) {
      std::vector<int64_t> cost(big.size());
      for (size_t i = 0; i < big.size(); ++i) {
        const FieldSpec& f = fields[big[i]];
        cost[i] = f.ny * f.nx * std::max<int64_t>(f.iterations, 1);
      }
      lane_of = rlic::SideWork::assign(cost, static_cast<int>(side.streams.size()));
    }
    for (size_t i = 0; i < big.size() && rc == RLIC_OK; ++i) {
      const int64_t k = big[i];
      const FieldSpec& f = fields[k];
      const hipStream_t s = side.streams.empty() ? stream : side.streams[lane_of[i]];
      const rlic::FieldGeom g{static_cast<int32_t>(f.ny), static_cast<int32_t>(f.nx),
                              f.x_left, f.x_right, f.y_left, f.y_right};
      rc = run_on_device<T>(f.texture, f.tex_sy, f.tex_sx, f.u, f.u_sy, f.u_sx, f.v, f.v_sy,
                            f.v_sx, static_cast<int32_t>(f.ny), static_cast<int32_t>(f.nx),
                            f.kernel, static_cast<int32_t>(f.kernel_len), f.uv_mode, g,
                            f.iterations, f.out, profile, s);
      if (rc != RLIC_OK) {
        t_last_error = "field " + std::to_string(k) + ": " + t_last_error;
        side.finish();
        (void)hipStreamSynchronize(stream);
        return rc;
      }
    }
    if (!rest.empty()) {
      rc = rlic::batch_run<T, FieldSpec>(rest.data(), static_cast<int64_t>(rest.size()), profile,
                                         stream, scratch, err);
    }
    // the caller's stream (and the batch timer) end after the side work
    if (hipError_t e = side.join(); e != hipSuccess && rc == RLIC_OK) {
      rc = RLIC_ERR_HIP;
      err = std::string("batched call: ") + hipGetErrorString(e);
    }
    if (rc == RLIC_OK) {
      if (hipError_t e = hipStreamSynchronize(stream); e != hipSuccess) {
        (void)hipGetLastError();
        rc = RLIC_ERR_HIP;
        err = std::string("batched call: ") + hipGetErrorString(e);
      }
    }
  }
  if (rc != RLIC_OK) {
    (void)hipStreamSynchronize(stream);
    return set_error(rc, "%s", err.c_str());
  }
  return RLIC_OK;
}

template <typename T>
int convolve_entry(const T* tex, int64_t tsy, int64_t tsx, const T* u, int64_t usy, int64_t usx,
                   const T* v, int64_t vsy, int64_t vsx, int64_t ny, int64_t nx, const T* kernel,
                   int64_t klen, int uv_mode, int xl, int xr, int yl, int yr, int64_t iterations,
                   T* out, int location, int profile, void* stream_ptr) {
  t_last_error.clear();
  if (ny < 0 || nx < 0) return set_error(RLIC_ERR_INVALID_ARGUMENT, "negative image size");
  if (klen < 0) return set_error(RLIC_ERR_INVALID_ARGUMENT, "negative kernel length");
  if (iterations < 0)
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "Invalid number of iterations: %lld",
                     static_cast<long long>(iterations));
  if (uv_mode != RLIC_UV_VELOCITY && uv_mode != RLIC_UV_POLARIZATION)
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "unknown uv_mode %d", uv_mode);
  for (int b : {xl, xr, yl, yr}) {
    if (b != RLIC_BOUND_CLOSED && b != RLIC_BOUND_PERIODIC)
      return set_error(RLIC_ERR_INVALID_ARGUMENT, "unknown boundary %d", b);
  }
  if (location != RLIC_HOST && location != RLIC_DEVICE)
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "unknown memory location %d", location);
  if (profile < 0 || profile > 3)
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "unknown arithmetic profile %d", profile);
  if (ny == 0 || nx == 0) return RLIC_OK;  // nothing to compute (the reference loops zero times)
  if (const int rc = check_size(ny, nx, klen); rc != RLIC_OK) return rc;
  const int64_t n = ny * nx;
  if (!tex || !u || !v || !out)
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "null array pointer");
  if (iterations > 0 && klen == 0)
    return set_error(RLIC_ERR_EMPTY_KERNEL,
                     "kernel is empty: kernel[len(kernel) // 2] does not exist");
  if (klen > 0 && !kernel) return set_error(RLIC_ERR_INVALID_ARGUMENT, "null kernel pointer");

  int dev = 0;
  RLIC_HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
  configure_pool(dev);
  hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
  const rlic::FieldGeom g{static_cast<int32_t>(ny), static_cast<int32_t>(nx), xl, xr, yl, yr};

  if (location == RLIC_DEVICE) {
    return run_on_device<T>(tex, tsy, tsx, u, usy, usx, v, vsy, vsx, static_cast<int32_t>(ny),
                            static_cast<int32_t>(nx), kernel, static_cast<int32_t>(klen),
                            uv_mode, g, iterations, out, profile, stream);
  }

  // Host arrays: upload the memory extent of each strided view, run, download.
  DevMem d_tex, d_u, d_v, d_k, d_out;
  const T* dev_in[3];
  const T* host_in[3] = {tex, u, v};
  const int64_t strides[3][2] = {{tsy, tsx}, {usy, usx}, {vsy, vsx}};
  DevMem* bufs[3] = {&d_tex, &d_u, &d_v};
  for (int a = 0; a < 3; ++a) {
    int64_t lo, hi;
    view_span(ny, nx, strides[a][0], strides[a][1], lo, hi);
    const size_t count = static_cast<size_t>(hi - lo + 1);
    RLIC_HIP_TRY(bufs[a]->alloc(sizeof(T) * count, stream), "allocating input");
    RLIC_HIP_TRY(hipMemcpyAsync(bufs[a]->ptr, host_in[a] + lo, sizeof(T) * count,
                                hipMemcpyHostToDevice, stream),
                 "uploading input");
    dev_in[a] = static_cast<const T*>(bufs[a]->ptr) - lo;
  }
  RLIC_HIP_TRY(d_k.alloc(sizeof(T) * (klen > 0 ? klen : 1), stream), "allocating kernel");
  if (klen > 0) {
    RLIC_HIP_TRY(hipMemcpyAsync(d_k.ptr, kernel, sizeof(T) * klen, hipMemcpyHostToDevice, stream),
                 "uploading kernel");
  }
  RLIC_HIP_TRY(d_out.alloc(sizeof(T) * n, stream), "allocating output");
  const int rc = run_on_device<T>(dev_in[0], tsy, tsx, dev_in[1], usy, usx, dev_in[2], vsy, vsx,
                                  static_cast<int32_t>(ny), static_cast<int32_t>(nx),
                                  static_cast<const T*>(d_k.ptr), static_cast<int32_t>(klen),
                                  uv_mode, g, iterations, static_cast<T*>(d_out.ptr), profile,
                                  stream);
  if (rc != RLIC_OK) {
    (void)hipStreamSynchronize(stream);
    return rc;
  }
  RLIC_HIP_TRY(hipMemcpyAsync(out, d_out.ptr, sizeof(T) * n, hipMemcpyDeviceToHost, stream),
               "downloading output");
  RLIC_HIP_TRY(hipStreamSynchronize(stream), "synchronizing");
  return RLIC_OK;
}

template <typename T, typename FieldSpec>
int server_admit(rlic::Server* server, const FieldSpec* f, int keep, int64_t* handle) {
  t_last_error.clear();
  if (!server || !f || !handle) return set_error(RLIC_ERR_INVALID_ARGUMENT, "null argument");
  const int rc = validate<T>(f->texture, f->u, f->v, f->ny, f->nx, f->kernel, f->kernel_len,
                             f->uv_mode, f->x_left, f->x_right, f->y_left, f->y_right,
                             f->iterations, f->texture);
  if (rc != RLIC_OK) return rc;
  if (f->iterations < 1)
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "a served field needs at least one pass");
  std::string err;
  const bool record = keep != 0 || f->iterations > 1;
  // an empty image is held as 0 x 0 with no weights (its output is empty
  // whatever its sides and kernel)
  const bool empty = f->ny == 0 || f->nx == 0;
  const int st = server->admit<T>(f->texture, f->tex_sy, f->tex_sx, f->u, f->u_sy, f->u_sx,
                                  f->v, f->v_sy, f->v_sx, empty ? 0 : f->ny, empty ? 0 : f->nx,
                                  f->kernel, empty ? 0 : f->kernel_len, f->uv_mode, f->x_left,
                                  f->x_right, f->y_left, f->y_right, f->iterations, record,
                                  handle, err);
  return st == RLIC_OK ? st : set_error(st, "%s", err.c_str());
}

template <typename T>
int64_t server_field_bytes(int64_t ny, int64_t nx, int64_t klen, int keep, int64_t iterations) {
  if (ny < 0 || nx < 0 || klen < 0) return -1;
  if (ny == 0 || nx == 0) ny = nx = klen = 0;  // held as 0 x 0 (see server_admit)
  else if (check_size(ny, nx, klen) != RLIC_OK) return -1;
  return static_cast<int64_t>(
      rlic::Server::field_bytes<T>(ny, nx, klen, keep != 0 || iterations > 1));
}

// Allocation-free convolve: the same checks and computation as
// convolve_entry with device arrays, its scratch in the fixed device array
// (see rlic_convolve_noalloc_* below).
int noalloc_scratch_ptr(void** ptr);

template <typename T>
int convolve_noalloc_entry(const T* tex, int64_t tsy, int64_t tsx, const T* u, int64_t usy,
                           int64_t usx, const T* v, int64_t vsy, int64_t vsx, int64_t ny,
                           int64_t nx, const T* kernel, int64_t klen, int uv_mode, int xl, int xr,
                           int yl, int yr, int64_t iterations, T* out, int profile,
                           void* stream_ptr) {
  t_last_error.clear();
  if (profile < 0 || profile > 3)
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "unknown arithmetic profile %d", profile);
  if (const int rc = validate<T>(tex, u, v, ny, nx, kernel, klen, uv_mode, xl, xr, yl, yr,
                                 iterations, out);
      rc != RLIC_OK)
    return rc;
  if (ny == 0 || nx == 0) return RLIC_OK;  // nothing to compute
  if (iterations > 0) {
    // weights, cells and ping-pong buffer must fit; the move record may not
    const size_t need = scratch_layout<T>(static_cast<int32_t>(ny), static_cast<int32_t>(nx),
                                          static_cast<int32_t>(klen), iterations > 1)
                            .codes;
    if (need > static_cast<size_t>(RLIC_NOALLOC_SCRATCH_BYTES))
      return set_error(RLIC_ERR_INVALID_ARGUMENT,
                       "an image of %lld x %lld pixels with a kernel of %lld weights needs %zu "
                       "bytes of scratch, more than the %lld bytes of the allocation-free "
                       "entry point (use rlic_convolve_f32/_f64)",
                       static_cast<long long>(ny), static_cast<long long>(nx),
                       static_cast<long long>(klen), need,
                       static_cast<long long>(RLIC_NOALLOC_SCRATCH_BYTES));
  }
  void* ptr = nullptr;
  if (iterations > 0) {
    if (const int rc = noalloc_scratch_ptr(&ptr); rc != RLIC_OK) return rc;
  }
  struct Guard {  // no HIP objects are created for this call's launches
    Guard() { t_noalloc = true; }
    ~Guard() { t_noalloc = false; }
  } guard;
  const FixedScratch fixed{ptr, static_cast<size_t>(RLIC_NOALLOC_SCRATCH_BYTES)};
  const rlic::FieldGeom g{static_cast<int32_t>(ny), static_cast<int32_t>(nx), xl, xr, yl, yr};
  return run_on_device<T>(tex, tsy, tsx, u, usy, usx, v, vsy, vsx, static_cast<int32_t>(ny),
                          static_cast<int32_t>(nx), kernel, static_cast<int32_t>(klen), uv_mode,
                          g, iterations, out, profile, static_cast<hipStream_t>(stream_ptr),
                          &fixed);
}

// Largest bin count of an equalization (the counters are 4 bytes each).
constexpr int64_t kMaxBins = int64_t{1} << 24;

// Largest counter array of a tile-interpolated equalization (tiles x nbins
// 4-byte counters).
constexpr int64_t kMaxTileCounters = int64_t{1} << 30;

// method 0: global; rlic::kEqSlidingTile / kEqTileInterpolation with a
// tile_h x tile_w window or tile.
template <typename T>
int equalize_entry(const T* image, int64_t sy, int64_t sx, int64_t ny, int64_t nx, int64_t nbins,
                   T* out, int location, void* stream_ptr, int method = 0, int64_t tile_h = 1,
                   int64_t tile_w = 1) {
  t_last_error.clear();
  if (ny < 0 || nx < 0) return set_error(RLIC_ERR_INVALID_ARGUMENT, "negative image size");
  if (nbins < 1 || nbins > kMaxBins)
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "nbins must be in [1, %lld], got %lld",
                     static_cast<long long>(kMaxBins), static_cast<long long>(nbins));
  if (method != 0 && (tile_h < 1 || tile_w < 1 || tile_h > INT32_MAX || tile_w > INT32_MAX))
    return set_error(RLIC_ERR_INVALID_ARGUMENT,
                     "tile_shape must be positive and below 2^31, got (%lld, %lld)",
                     static_cast<long long>(tile_h), static_cast<long long>(tile_w));
  if (location != RLIC_HOST && location != RLIC_DEVICE)
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "unknown memory location %d", location);
  if (ny == 0 || nx == 0) return RLIC_OK;  // an empty image: nothing to do
  if (ny > INT32_MAX || nx > INT32_MAX || ny * nx > INT32_MAX)
    return set_error(RLIC_ERR_INVALID_ARGUMENT,
                     "image of %lld x %lld pixels exceeds the supported 2^31-1 pixels",
                     static_cast<long long>(ny), static_cast<long long>(nx));
  if (method == rlic::kEqTileInterpolation) {
    const int64_t tiles = ((ny + std::min(tile_h, ny) - 1) / std::min(tile_h, ny)) *
                          ((nx + std::min(tile_w, nx) - 1) / std::min(tile_w, nx));
    if (tiles > kMaxTileCounters / nbins)
      return set_error(RLIC_ERR_INVALID_ARGUMENT,
                       "%lld tiles x %lld bins exceed the %lld tile counters supported",
                       static_cast<long long>(tiles), static_cast<long long>(nbins),
                       static_cast<long long>(kMaxTileCounters));
  }
  if (!image || !out) return set_error(RLIC_ERR_INVALID_ARGUMENT, "null array pointer");
  int dev = 0;
  RLIC_HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
  configure_pool(dev);
  hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
  const int64_t n = ny * nx;
  const int32_t bins = static_cast<int32_t>(nbins);
  auto run = [&](const T* img, int64_t isy, int64_t isx, T* dst) -> int {
    rlic::CallScratch scratch;
    const int32_t h = static_cast<int32_t>(ny), w = static_cast<int32_t>(nx);
    const int32_t th = static_cast<int32_t>(tile_h), tw = static_cast<int32_t>(tile_w);
    if (method == 0) {
      RLIC_HIP_TRY(scratch.get(rlic::equalize_scratch_bytes<T>(bins), stream),
                   "allocating equalization scratch");
      LaunchTimer timer("lic_equalize", stream);
      RLIC_HIP_TRY(rlic::launch_equalize<T>(img, isy, isx, h, w, bins, dst, scratch.ptr(), stream),
                   "equalize");
      return RLIC_OK;
    }
    RLIC_HIP_TRY(scratch.get(rlic::equalize_tiles_scratch_bytes<T>(method, h, w, bins, th, tw),
                             stream),
                 "allocating equalization scratch");
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) {
      (void)hipGetLastError();
      lds = 65536;
    }
    LaunchTimer timer(method == rlic::kEqSlidingTile ? "lic_equalize_sliding_tile"
                                                     : "lic_equalize_tile_interpolation",
                      stream);
    RLIC_HIP_TRY(rlic::launch_equalize_tiles<T>(method, img, isy, isx, h, w, bins, th, tw, dst,
                                                scratch.ptr(), lds, stream),
                 "equalize");
    return RLIC_OK;
  };
  if (location == RLIC_DEVICE) return run(image, sy, sx, out);
  // host arrays: upload the memory extent of the strided view, run, download
  DevMem d_in, d_out;
  int64_t lo, hi;
  view_span(ny, nx, sy, sx, lo, hi);
  const size_t count = static_cast<size_t>(hi - lo + 1);
  RLIC_HIP_TRY(d_in.alloc(sizeof(T) * count, stream), "allocating input");
  RLIC_HIP_TRY(hipMemcpyAsync(d_in.ptr, image + lo, sizeof(T) * count, hipMemcpyHostToDevice,
                              stream),
               "uploading input");
  RLIC_HIP_TRY(d_out.alloc(sizeof(T) * static_cast<size_t>(n), stream), "allocating output");
  const int rc = run(static_cast<const T*>(d_in.ptr) - lo, sy, sx, static_cast<T*>(d_out.ptr));
  if (rc != RLIC_OK) {
    (void)hipStreamSynchronize(stream);
    return rc;
  }
  RLIC_HIP_TRY(hipMemcpyAsync(out, d_out.ptr, sizeof(T) * static_cast<size_t>(n),
                              hipMemcpyDeviceToHost, stream),
               "downloading output");
  RLIC_HIP_TRY(hipStreamSynchronize(stream), "synchronizing");
  return RLIC_OK;
}

template <typename T>
int server_read(rlic::Server* server, int64_t handle, T* out) {
  t_last_error.clear();
  if (!server) return set_error(RLIC_ERR_INVALID_ARGUMENT, "null server");
  std::string err;
  const int st = server->read<T>(handle, out, err);
  return st == RLIC_OK ? st : set_error(st, "%s", err.c_str());
}

}  // namespace

namespace rlic {

template <typename T>
hipError_t replay_pass(const T* src, T* dst, const T* kernel, int32_t klen, const FieldGeom& g,
                       int profile, const uint4* codes, hipStream_t stream) {
  const ReplayPlan plan = default_plan<T>(klen, static_cast<int64_t>(g.ny) * g.nx);
  LaunchTimer timer(plan_kernel_name(plan), stream);
  return launch_plan<T>(plan, src, dst, kernel, klen, g, profile, codes, stream);
}

template hipError_t replay_pass<float>(const float*, float*, const float*, int32_t,
                                       const FieldGeom&, int, const uint4*, hipStream_t);
template hipError_t replay_pass<double>(const double*, double*, const double*, int32_t,
                                        const FieldGeom&, int, const uint4*, hipStream_t);

}  // namespace rlic

namespace {

template <typename T>
int fill_noise(T* out, int64_t ny, int64_t nx, uint64_t seed, void* stream) {
  t_last_error.clear();
  if (ny < 0 || nx < 0) return set_error(RLIC_ERR_INVALID_ARGUMENT, "negative image size");
  if (ny > 0 && nx > INT64_MAX / ny) return set_error(RLIC_ERR_INVALID_ARGUMENT, "image too large");
  if (ny * nx > 0 && !out) return set_error(RLIC_ERR_INVALID_ARGUMENT, "null array pointer");
  RLIC_HIP_TRY(rlic::launch_fill_noise<T>(out, ny * nx, seed, static_cast<hipStream_t>(stream)),
               "noise");
  return RLIC_OK;
}

template <typename T>
int fill_modes(T* out, int64_t ny, int64_t nx, int32_t n_modes, const double* amp,
               const int32_t* kx, const int32_t* ky, const double* phase, void* stream) {
  t_last_error.clear();
  if (ny < 0 || nx < 0) return set_error(RLIC_ERR_INVALID_ARGUMENT, "negative image size");
  if (ny > INT32_MAX / 2 || nx > INT32_MAX / 2)
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "image too large");
  if (n_modes < 0 || n_modes > rlic::kMaxModes)
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "n_modes must be in [0, %d]", rlic::kMaxModes);
  if (n_modes > 0 && (!amp || !kx || !ky || !phase))
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "null mode array");
  if (ny * nx > 0 && !out) return set_error(RLIC_ERR_INVALID_ARGUMENT, "null array pointer");
  rlic::ModeSet modes{};
  modes.n = n_modes;
  for (int32_t m = 0; m < n_modes; ++m) {
    modes.amp[m] = amp[m];
    modes.phase[m] = phase[m];
    modes.kx[m] = kx[m];
    modes.ky[m] = ky[m];
  }
  RLIC_HIP_TRY(rlic::launch_fill_modes<T>(out, static_cast<int32_t>(ny), static_cast<int32_t>(nx),
                                          modes, static_cast<hipStream_t>(stream)),
               "Fourier modes");
  return RLIC_OK;
}

}  // namespace

// The allocation-free entry points' scratch: one fixed array per device,
// held until the process ends.  In host code `scratch` names a host-side
// stand-in, not device memory: host code uses only the device address that
// hipGetSymbolAddress returns for the current device, taken once per
// device on that device's first allocation-free call and kept.
static __device__ __align__(256) unsigned char scratch[RLIC_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 256;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices] = {};

// The current device's address of `scratch` (looked up on its first use).
int noalloc_scratch_ptr(void** ptr) {
  int dev = -1;
  RLIC_HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices)
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_error(e, "locating the fixed scratch array");
    if (!p) return set_error(RLIC_ERR_HIP, "locating the fixed scratch array: null address");
    g_noalloc_ptr[dev] = p;
  }
  *ptr = g_noalloc_ptr[dev];
  return RLIC_OK;
}

}  // namespace

extern "C" {

int rlic_convolve_f32(const float* texture, int64_t tex_sy, int64_t tex_sx, const float* u,
                      int64_t u_sy, int64_t u_sx, const float* v, int64_t v_sy, int64_t v_sx,
                      int64_t ny, int64_t nx, const float* kernel, int64_t kernel_len,
                      int uv_mode, int x_left, int x_right, int y_left, int y_right,
                      int64_t iterations, float* out, int location, int profile, void* stream) {
  return convolve_entry<float>(texture, tex_sy, tex_sx, u, u_sy, u_sx, v, v_sy, v_sx, ny, nx,
                               kernel, kernel_len, uv_mode, x_left, x_right, y_left, y_right,
                               iterations, out, location, profile, stream);
}

int rlic_convolve_f64(const double* texture, int64_t tex_sy, int64_t tex_sx, const double* u,
                      int64_t u_sy, int64_t u_sx, const double* v, int64_t v_sy, int64_t v_sx,
                      int64_t ny, int64_t nx, const double* kernel, int64_t kernel_len,
                      int uv_mode, int x_left, int x_right, int y_left, int y_right,
                      int64_t iterations, double* out, int location, int profile,
                      void* stream) {
  return convolve_entry<double>(texture, tex_sy, tex_sx, u, u_sy, u_sx, v, v_sy, v_sx, ny, nx,
                                kernel, kernel_len, uv_mode, x_left, x_right, y_left, y_right,
                                iterations, out, location, profile, stream);
}

int rlic_convolve_noalloc_f32(const float* texture, int64_t tex_sy, int64_t tex_sx,
                              const float* u, int64_t u_sy, int64_t u_sx, const float* v,
                              int64_t v_sy, int64_t v_sx, int64_t ny, int64_t nx,
                              const float* kernel, int64_t kernel_len, int uv_mode, int x_left,
                              int x_right, int y_left, int y_right, int64_t iterations,
                              float* out, int profile, void* stream) {
  return convolve_noalloc_entry<float>(texture, tex_sy, tex_sx, u, u_sy, u_sx, v, v_sy, v_sx,
                                       ny, nx, kernel, kernel_len, uv_mode, x_left, x_right,
                                       y_left, y_right, iterations, out, profile, stream);
}

int rlic_convolve_noalloc_f64(const double* texture, int64_t tex_sy, int64_t tex_sx,
                              const double* u, int64_t u_sy, int64_t u_sx, const double* v,
                              int64_t v_sy, int64_t v_sx, int64_t ny, int64_t nx,
                              const double* kernel, int64_t kernel_len, int uv_mode, int x_left,
                              int x_right, int y_left, int y_right, int64_t iterations,
                              double* out, int profile, void* stream) {
  return convolve_noalloc_entry<double>(texture, tex_sy, tex_sx, u, u_sy, u_sx, v, v_sy, v_sx,
                                        ny, nx, kernel, kernel_len, uv_mode, x_left, x_right,
                                        y_left, y_right, iterations, out, profile, stream);
}

int rlic_equalize_histogram_f32(const float* image, int64_t sy, int64_t sx, int64_t ny,
                                int64_t nx, int64_t nbins, float* out, int location,
                                void* stream) {
  return equalize_entry<float>(image, sy, sx, ny, nx, nbins, out, location, stream);
}

int rlic_equalize_histogram_f64(const double* image, int64_t sy, int64_t sx, int64_t ny,
                                int64_t nx, int64_t nbins, double* out, int location,
                                void* stream) {
  return equalize_entry<double>(image, sy, sx, ny, nx, nbins, out, location, stream);
}

int rlic_equalize_histogram_sliding_tile_f32(const float* image, int64_t sy, int64_t sx,
                                             int64_t ny, int64_t nx, int64_t nbins,
                                             int64_t tile_h, int64_t tile_w, float* out,
                                             int location, void* stream) {
  return equalize_entry<float>(image, sy, sx, ny, nx, nbins, out, location, stream,
                               rlic::kEqSlidingTile, tile_h, tile_w);
}

int rlic_equalize_histogram_sliding_tile_f64(const double* image, int64_t sy, int64_t sx,
                                             int64_t ny, int64_t nx, int64_t nbins,
                                             int64_t tile_h, int64_t tile_w, double* out,
                                             int location, void* stream) {
  return equalize_entry<double>(image, sy, sx, ny, nx, nbins, out, location, stream,
                                rlic::kEqSlidingTile, tile_h, tile_w);
}

int rlic_equalize_histogram_tile_interpolation_f32(const float* image, int64_t sy, int64_t sx,
                                                   int64_t ny, int64_t nx, int64_t nbins,
                                                   int64_t tile_h, int64_t tile_w, float* out,
                                                   int location, void* stream) {
  return equalize_entry<float>(image, sy, sx, ny, nx, nbins, out, location, stream,
                               rlic::kEqTileInterpolation, tile_h, tile_w);
}

int rlic_equalize_histogram_tile_interpolation_f64(const double* image, int64_t sy, int64_t sx,
                                                   int64_t ny, int64_t nx, int64_t nbins,
                                                   int64_t tile_h, int64_t tile_w, double* out,
                                                   int location, void* stream) {
  return equalize_entry<double>(image, sy, sx, ny, nx, nbins, out, location, stream,
                                rlic::kEqTileInterpolation, tile_h, tile_w);
}

int rlic_convolve_batch_f32(const rlic_field_f32* fields, int64_t n_fields, int profile,
                            void* stream) {
  return batch_entry<float, rlic_field_f32>(fields, n_fields, profile, stream);
}

int rlic_convolve_batch_f64(const rlic_field_f64* fields, int64_t n_fields, int profile,
                            void* stream) {
  return batch_entry<double, rlic_field_f64>(fields, n_fields, profile, stream);
}

int rlic_batch_check_values_f32(const rlic_field_f32* fields, int64_t n_fields,
                                int64_t* first_bad, int* what, void* stream) {
  return batch_check_entry<float, rlic_field_f32>(fields, n_fields, first_bad, what, stream);
}

int rlic_batch_check_values_f64(const rlic_field_f64* fields, int64_t n_fields,
                                int64_t* first_bad, int* what, void* stream) {
  return batch_check_entry<double, rlic_field_f64>(fields, n_fields, first_bad, what, stream);
}

int rlic_check_values_f32(const float* texture, int64_t ny, int64_t nx, int64_t tex_sy,
                          int64_t tex_sx, const float* kernel, int64_t kernel_len, int* what,
                          void* stream) {
  return check_entry<float>(texture, ny, nx, tex_sy, tex_sx, kernel, kernel_len, what, stream);
}

int rlic_check_values_f64(const double* texture, int64_t ny, int64_t nx, int64_t tex_sy,
                          int64_t tex_sx, const double* kernel, int64_t kernel_len, int* what,
                          void* stream) {
  return check_entry<double>(texture, ny, nx, tex_sy, tex_sx, kernel, kernel_len, what, stream);
}

struct rlic_server {
  rlic::Server core;
  rlic_server(int64_t budget, int profile, hipStream_t stream) : core(budget, profile, stream) {}
};

rlic_server* rlic_server_create(int64_t budget_bytes, int profile, void* stream) {
  t_last_error.clear();
  if (budget_bytes < 0) {
    set_error(RLIC_ERR_INVALID_ARGUMENT, "negative budget");
    return nullptr;
  }
  if (profile < 0 || profile > 3) {
    set_error(RLIC_ERR_INVALID_ARGUMENT, "unknown arithmetic profile %d", profile);
    return nullptr;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    hip_error(hipGetLastError(), "hipGetDevice");
    return nullptr;
  }
  configure_pool(dev);
  return new rlic_server(budget_bytes, profile, static_cast<hipStream_t>(stream));
}

void rlic_server_destroy(rlic_server* server) { delete server; }

int rlic_server_admit_f32(rlic_server* server, const rlic_field_f32* field, int keep,
                          int64_t* handle) {
  return server_admit<float>(server ? &server->core : nullptr, field, keep, handle);
}

int rlic_server_admit_f64(rlic_server* server, const rlic_field_f64* field, int keep,
                          int64_t* handle) {
  return server_admit<double>(server ? &server->core : nullptr, field, keep, handle);
}

int64_t rlic_server_field_bytes_f32(int64_t ny, int64_t nx, int64_t kernel_len, int keep,
                                    int64_t iterations) {
  return server_field_bytes<float>(ny, nx, kernel_len, keep, iterations);
}

int64_t rlic_server_field_bytes_f64(int64_t ny, int64_t nx, int64_t kernel_len, int keep,
                                    int64_t iterations) {
  return server_field_bytes<double>(ny, nx, kernel_len, keep, iterations);
}

int rlic_server_step(rlic_server* server, int64_t* finished, int64_t capacity,
                     int64_t* n_finished) {
  t_last_error.clear();
  if (!server || !n_finished || capacity < 0 || (capacity > 0 && !finished))
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "invalid argument");
  int64_t active = 0;
  server->core.usage(nullptr, nullptr, &active);
  if (capacity < active)
    return set_error(RLIC_ERR_INVALID_ARGUMENT,
                     "capacity %lld is below the %lld active fields",
                     static_cast<long long>(capacity), static_cast<long long>(active));
  std::vector<int64_t> done;
  std::string err;
  int st;
  {
    LaunchTimer timer("lic_serve_step", server->core.stream());
    st = server->core.step(done, err);
  }
  if (st != RLIC_OK) return set_error(st, "%s", err.c_str());
  std::copy(done.begin(), done.end(), finished);
  *n_finished = static_cast<int64_t>(done.size());
  return RLIC_OK;
}

int rlic_server_extend(rlic_server* server, int64_t handle, int64_t more) {
  t_last_error.clear();
  if (!server || more < 0) return set_error(RLIC_ERR_INVALID_ARGUMENT, "invalid argument");
  std::string err;
  const int st = server->core.extend(handle, more, err);
  return st == RLIC_OK ? st : set_error(st, "%s", err.c_str());
}

int rlic_server_read_f32(rlic_server* server, int64_t handle, float* out) {
  return server_read<float>(server ? &server->core : nullptr, handle, out);
}

int rlic_server_read_f64(rlic_server* server, int64_t handle, double* out) {
  return server_read<double>(server ? &server->core : nullptr, handle, out);
}

int rlic_server_release(rlic_server* server, int64_t handle) {
  t_last_error.clear();
  if (!server) return set_error(RLIC_ERR_INVALID_ARGUMENT, "null server");
  std::string err;
  const int st = server->core.release(handle, err);
  return st == RLIC_OK ? st : set_error(st, "%s", err.c_str());
}

int rlic_server_usage(rlic_server* server, int64_t* used_bytes, int64_t* resident,
                      int64_t* active) {
  t_last_error.clear();
  if (!server) return set_error(RLIC_ERR_INVALID_ARGUMENT, "null server");
  server->core.usage(used_bytes, resident, active);
  return RLIC_OK;
}

const char* rlic_last_error(void) { return t_last_error.c_str(); }

int rlic_profiler_enable(int on) {
  std::lock_guard<std::mutex> lock(g_prof_mu);
  g_prof_on = on != 0;
  return RLIC_OK;
}

int64_t rlic_profiler_read(char* buf, int64_t len) {
  std::lock_guard<std::mutex> lock(g_prof_mu);
  for (const TimedLaunch& t : g_prof_pending) {
    float ms = 0.f;
    if (hipEventSynchronize(t.stop) == hipSuccess &&
        hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess) {
      auto& agg = g_prof_done[t.name];
      agg.first += t.units;
      agg.second += ms;
    }
    (void)hipEventDestroy(t.start);
    (void)hipEventDestroy(t.stop);
  }
  g_prof_pending.clear();
  std::string json = "{";
  bool first = true;
  for (const auto& kv : g_prof_done) {
    char item[256];
    snprintf(item, sizeof(item), "%s\"%s\": {\"count\": %lld, \"total_ms\": %.6f}",
             first ? "" : ", ", kv.first.c_str(), static_cast<long long>(kv.second.first),
             kv.second.second);
    json += item;
    first = false;
  }
  json += "}";
  const int64_t need = static_cast<int64_t>(json.size()) + 1;
  if (buf && len >= need) {
    std::memcpy(buf, json.c_str(), need);
    g_prof_done.clear();  // handed over: start a new record
  }
  return need;
}

int rlic_div_check_f32(const float* a, const float* b, float* q, int64_t n, void* stream) {
  t_last_error.clear();
  if (n < 0 || (n > 0 && (!a || !b || !q)))
    return set_error(RLIC_ERR_INVALID_ARGUMENT, "invalid argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  RLIC_HIP_TRY(rlic::launch_div_check(a, b, q, n, s), "division check");
  RLIC_HIP_TRY(hipStreamSynchronize(s), "division check");
  return RLIC_OK;
}

int rlic_fill_noise_f32(float* out, int64_t ny, int64_t nx, uint64_t seed, void* stream) {
  return fill_noise<float>(out, ny, nx, seed, stream);
}

int rlic_fill_noise_f64(double* out, int64_t ny, int64_t nx, uint64_t seed, void* stream) {
  return fill_noise<double>(out, ny, nx, seed, stream);
}

int rlic_fill_modes_f32(float* out, int64_t ny, int64_t nx, int32_t n_modes, const double* amp,
                        const int32_t* kx, const int32_t* ky, const double* phase, void* stream) {
  return fill_modes<float>(out, ny, nx, n_modes, amp, kx, ky, phase, stream);
}

int rlic_fill_modes_f64(double* out, int64_t ny, int64_t nx, int32_t n_modes, const double* amp,
                        const int32_t* kx, const int32_t* ky, const double* phase, void* stream) {
  return fill_modes<double>(out, ny, nx, n_modes, amp, kx, ky, phase, stream);
}

int rlic_scratch_release(void) {
  t_last_error.clear();
  RLIC_HIP_TRY(rlic::release_cached_scratch(), "releasing cached scratch");
  return RLIC_OK;
}

int64_t rlic_autotune_read(char* buf, int64_t len) {
  std::lock_guard<std::mutex> lock(g_tuner.mu);
  tuner_load_locked();
  std::string json = "[";
  bool first = true;
  for (const auto& kv : g_tuner.table) {
    const TuneKey& k = kv.first;
    const TuneEntry& e = kv.second;
    char item[320];
    snprintf(item, sizeof(item),
             "%s{\"dtype_size\": %d, \"ny\": %d, \"nx\": %d, \"kernel_len\": %d, \"profile\": %d, "
             "\"replay\": \"%s\", \"tile\": [%d, %d], \"band\": %d, \"wave_w_log2\": %d, "
             "\"pass_ms\": %.6f}",
             first ? "" : ", ", k.dtype_size, k.ny, k.nx, k.klen, k.profile,
             e.plan.lds == 2 ? "lds_hybrid" : (e.plan.lds ? "lds" : "l2"), e.plan.th, e.plan.tw,
             e.plan.band, e.plan.wave_w_log2,
             e.ms);
    json += item;
    first = false;
  }
  json += "]";
  const int64_t need = static_cast<int64_t>(json.size()) + 1;
  if (buf && len >= need) std::memcpy(buf, json.c_str(), need);
  return need;
}

int64_t rlic_autotune_load(const char* path) {
  t_last_error.clear();
  if (!path) return set_error(RLIC_ERR_INVALID_ARGUMENT, "null path");
  std::lock_guard<std::mutex> lock(g_tuner.mu);
  const int64_t n = tuner_read_file_locked(path);
  if (n < 0) return set_error(RLIC_ERR_INVALID_ARGUMENT, "cannot open %s", path);
  return n;
}

int rlic_autotune_clear(void) {
  std::lock_guard<std::mutex> lock(g_tuner.mu);
  g_tuner.table.clear();
  g_tuner.loaded = true;  // a cleared table does not reload the cache file
  return RLIC_OK;
}

int rlic_trace_stats_read(int64_t* live_lane_steps, int64_t* executed_lane_steps) {
  std::lock_guard<std::mutex> lock(g_stats_mu);
  if (live_lane_steps) *live_lane_steps = g_stats_useful;
  if (executed_lane_steps) *executed_lane_steps = g_stats_executed;
  g_stats_useful = g_stats_executed = 0;
  return RLIC_OK;
}

const char* rlic_version(void) { return "0.1.0"; }

}  // extern "C"
