// Many fields in one persistent launch from a device work queue.
//
// Each field is one `rlic.convolve` call (src/lib.rs:408-443) with its own
// size, kernel, boundaries, uv_mode and iteration count.  Work items are
// square pixel regions of one pass of one field; a segment is the run of
// items of one (field, pass).  The waves of a persistent grid pop items from
// one atomic counter.  Pass 0 traces the streamlines from the packed cells
// (and records the moves when more passes follow); later passes replay the
// moves with gathers through L2.
//
// The batched call queues every pass of every field in one launch,
// pass-major (every field's pass 0, then every field's pass 1, ...), and a
// pass-p item waits until the field's pass p-1 is complete.  The items it
// waits for were queued earlier, so they were popped by waves that are
// already running and whose own waits point further back: every wait ends
// (a wait longer than a time limit still ends, and makes the call fail).
// A serving step queues one pass of each active field, with no waits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/rlic_hip.h"
#include "lic_batch.h"
#include "lic_device.h"
#include "lic_launch.h"
#include "lic_scratch.h"
#include "lic_steps.h"

namespace rlic {

constexpr int32_t kPendSlots = 8;       // segments a wave holds unpublished counts for
constexpr uint32_t kPublishItems = 32;  // completed items per release of a wave

// Kernel weights of every field into its zero-padded slot (one block per field).
template <typename T>
__global__ __launch_bounds__(256) void lic_pad_batch(const BatchField<T>* fields,
                                                     const T* const* kernels) {
  const BatchField<T>& f = fields[blockIdx.x];
  // an empty image has no pixel to weight, and its kernel pointer was not
  // checked (the host accepts any kernel argument of an empty field)
  if (f.g.ny == 0 || f.g.nx == 0) return;
  const T* src = kernels[blockIdx.x];
  T* dst = const_cast<T*>(f.kernel) - kKernelPad;
  const int32_t total = f.klen + 2 * kKernelPad;
  for (int32_t t = threadIdx.x; t < total; t += 256) {
    const int32_t k = t - kKernelPad;
    dst[t] = (k >= 0 && k < f.klen) ? src[k] : T(0);
  }
}

// Pack every field's (texture, u, v) into its guarded cell array (ring of
// guard cells, see guarded_cell); a field with zero passes gets its texture
// copied to `out` instead.  Block b covers 64 x 4 cells of the guarded
// array of the field whose pack_block0 range contains b.
template <typename T>
__global__ __launch_bounds__(256) void lic_pack_batch(const BatchField<T>* fields, int32_t nfields) {
  const int64_t b = blockIdx.x;
  int32_t lo = 0, hi = nfields - 1;
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (fields[mid].pack_block0 <= b) lo = mid; else hi = mid - 1;
  }
  const BatchField<T>& f = fields[lo];
  const int64_t local = b - f.pack_block0;
  const int32_t bx = (f.g.nx + 2 + 63) / 64;
  const int32_t gj = static_cast<int32_t>(local % bx) * 64 + (threadIdx.x & 63);
  const int32_t gi = static_cast<int32_t>(local / bx) * 4 + (threadIdx.x >> 6);
  if (gi >= f.g.ny + 2 || gj >= f.g.nx + 2) return;
  if (f.iterations == 0) {
    const int32_t i = gi - 1, j = gj - 1;
    if (i >= 0 && i < f.g.ny && j >= 0 && j < f.g.nx)
      f.out[static_cast<int64_t>(i) * f.g.nx + j] = f.tex[i * f.tsy + j * f.tsx];
    return;
  }
  f.cells[static_cast<int64_t>(gi) * (f.g.nx + 2) + gj] =
      guarded_cell(f.tex, f.tsy, f.tsx, f.u, f.usy, f.usx, f.v, f.vsy, f.vsx, f.g, gi, gj);
}

__global__ void lic_zero_u32(uint32_t* p, int64_t n) {
  for (int64_t t = blockIdx.x * 256 + threadIdx.x; t < n; t += gridDim.x * 256) p[t] = 0u;
}

template <typename T, bool FMA, bool BRANCHLESS, int POL, int REC, bool BUF>
__device__ __forceinline__ void batch_trace_pixel(const BatchField<T>& f,
                                                  const T* __restrict__ kernel, T* dst,
                                                  int32_t i0, int32_t j0) {
  const CellGather<T, BUF> cell(f.cells, guarded_cells(f.g.ny, f.g.nx));
  const StepGeom sg = make_step_geom(f.g, cell.unit);
  const int32_t kmid = f.klen / 2;
  const int32_t n_fwd = f.klen - 1 - kmid, n_bwd = kmid;
  const int64_t np = static_cast<int64_t>(f.g.ny) * f.g.nx;
  const int64_t p = static_cast<int64_t>(i0) * f.g.nx + j0;
  const uint32_t row0 =
      static_cast<uint32_t>(i0 + 1) * sg.grow + static_cast<uint32_t>(j0 + 1) * cell.unit;
  const CellVal<T> c0 = cell(row0);
  T acc = macc<FMA>(kernel[kmid], c0.x, T(0));  // src/lib.rs:375-383
  acc = trace_direction<T, FMA, BRANCHLESS, POL, true, REC, BUF>(
      acc, c0.u, c0.v, row0, cell, kernel, kmid, n_fwd, sg, f.codes, np, p, 0);
  acc = trace_direction<T, FMA, BRANCHLESS, POL, false, REC, BUF>(
      acc, c0.u, c0.v, row0, cell, kernel, kmid, n_bwd, sg, f.codes, np, p,
      chunks_of(n_fwd));
  dst[p] = acc;
}

template <typename T, bool FMA, bool BUF>
__device__ __forceinline__ void batch_replay_pixel(const BatchField<T>& f,
                                                   const T* __restrict__ kernel, const T* src,
                                                   T* dst, const uint32_t* delta, int32_t i0,
                                                   int32_t j0) {
  const ValueGather<T, BUF> x(src, static_cast<int64_t>(f.g.ny) * f.g.nx);
  const int32_t kmid = f.klen / 2;
  const int32_t n_fwd = f.klen - 1 - kmid, n_bwd = kmid;
  const int64_t np = static_cast<int64_t>(f.g.ny) * f.g.nx;
  const int64_t idx0 = static_cast<int64_t>(i0) * f.g.nx + j0;
  const uint32_t pos0 = static_cast<uint32_t>(idx0) * x.unit;
  T acc = macc<FMA>(kernel[kmid], x(pos0), T(0));
  acc = replay_direction<T, FMA, true, BUF>(acc, pos0, x, kernel, kmid, n_fwd, delta, f.codes,
                                            np, idx0, 0);
  acc = replay_direction<T, FMA, false, BUF>(acc, pos0, x, kernel, kmid, n_bwd, delta, f.codes,
                                             np, idx0, chunks_of(n_fwd));
  dst[idx0] = acc;
}

// The region of a work item, (8 << sq_log2) pixels on a side, as squares of
// 8 x 8 lanes; (i_base, j_base) is the lane's pixel in the first square.
template <typename T, bool FMA, bool BRANCHLESS, int POL, int REC, bool BUF>
__device__ __forceinline__ void batch_trace_region(const BatchField<T>& f,
                                                   const T* __restrict__ kernel, T* dst,
                                                   int32_t i_base, int32_t j_base,
                                                   int32_t sq_log2) {
  const int32_t sq_mask = (1 << sq_log2) - 1;
#pragma unroll 1
  for (int32_t st = 0; st < (1 << (2 * sq_log2)); ++st) {
    const int32_t i0 = i_base + (st >> sq_log2) * 8, j0 = j_base + (st & sq_mask) * 8;
    if (i0 < f.g.ny && j0 < f.g.nx)
      batch_trace_pixel<T, FMA, BRANCHLESS, POL, REC, BUF>(f, kernel, dst, i0, j0);
  }
}

template <typename T, bool FMA, bool BUF>
__device__ __forceinline__ void batch_replay_region(const BatchField<T>& f,
                                                    const T* __restrict__ kernel, const T* src,
                                                    T* dst, const uint32_t* delta,
                                                    int32_t i_base, int32_t j_base,
                                                    int32_t sq_log2) {
  const int32_t sq_mask = (1 << sq_log2) - 1;
#pragma unroll 1
  for (int32_t st = 0; st < (1 << (2 * sq_log2)); ++st) {
    const int32_t i0 = i_base + (st >> sq_log2) * 8, j0 = j_base + (st & sq_mask) * 8;
    if (i0 < f.g.ny && j0 < f.g.nx)
      batch_replay_pixel<T, FMA, BUF>(f, kernel, src, dst, delta, i0, j0);
  }
}

// 64-bit value of lane 0 (the first active lane when the whole wave runs),
// as a wave-uniform scalar.
__device__ __forceinline__ int64_t lane0_value(uint64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(x));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(x >> 32));
  return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}

// Persistent kernel: every wave is an independent worker.  A wave pops one
// work item (a square region of one pass of one field), waits if needed
// (DEPS) for the segment it depends on, computes the region's pixels as
// squares of 8 x 8 lanes, then (DEPS) counts the item done.  Regions are
// 8 << sq_log2 pixels on a side (16, 32 or 64).  There is no block barrier
// in the loop, so waves of one block never wait for each other, and every
// branch that steers the loop is wave-uniform.
template <typename T, bool FMA, bool BRANCHLESS, bool BUF, bool DEPS>
__global__ __launch_bounds__(256) void lic_batch_persistent(const BatchField<T>* __restrict__ fields,
                                                            const Segment* __restrict__ segs,
                                                            const T* __restrict__ kpool,
                                                            int32_t nsegs,
                                                            int64_t total_items,
                                                            unsigned long long* queue,
                                                            uint32_t* done, uint32_t* failed,
                                                            int64_t wait_ticks, int32_t sq_log2) {
  __shared__ uint32_t delta_s[4][16];
  const int32_t lane = threadIdx.x & 63;
  uint32_t* delta = delta_s[threadIdx.x >> 6];
  const int32_t py = lane >> 3, px = lane & 7;
  // Completed items are published in batches: per-segment counts of up to
  // kPendSlots segments are held by the wave and published together behind
  // one release fence (the fence writes back L2, too costly per item) once
  // kPublishItems items are pending, when a new segment finds no free slot,
  // before the wave waits (so no wave waits while it holds unpublished
  // work) and when the queue is drained.
  int32_t pseg[kPendSlots] = {};
  uint32_t pcnt[kPendSlots] = {};
  int32_t npend = 0;
  uint32_t pitems = 0;
  int32_t fenced_prev = -1;  // segment whose completion the last acquire covered
  auto publish = [&]() {
    if (npend != 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __builtin_amdgcn_wave_barrier();
      if (lane == 0) {
#pragma unroll
        for (int32_t k = 0; k < kPendSlots; ++k)
          if (k < npend)
            __hip_atomic_fetch_add(&done[pseg[k]], pcnt[k], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
      }
      npend = 0;
      pitems = 0;
    }
  };
  auto note_done = [&](int32_t sgi) {
    bool found = false;
#pragma unroll
    for (int32_t k = 0; k < kPendSlots; ++k) {
      if (k < npend && pseg[k] == sgi) {
        pcnt[k] += 1u;
        found = true;
      }
    }
    if (!found) {
      if (npend == kPendSlots) publish();
#pragma unroll
      for (int32_t k = 0; k < kPendSlots; ++k) {
        if (k == npend) {
          pseg[k] = sgi;
          pcnt[k] = 1u;
        }
      }
      ++npend;
    }
    if (++pitems >= kPublishItems) publish();
  };
  for (;;) {
    uint64_t got = 0;
    if (lane == 0) got = atomicAdd(queue, 1ull);
    const int64_t item = lane0_value(got);
    if (item >= total_items) break;  // the queue is drained
    int32_t lo = 0, hi = nsegs - 1;
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (segs[mid].item0 <= item) lo = mid; else hi = mid - 1;
    }
    const Segment seg = segs[lo];
    const BatchField<T>& f = fields[seg.field];
    if (DEPS && seg.prev >= 0) {
      // every item of this field's previous pass must be complete; those
      // items were queued earlier, hence popped by running waves whose own
      // waits point further back
      auto complete = [&]() {
        return __builtin_amdgcn_readfirstlane(__hip_atomic_load(
                   &done[seg.prev], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >=
               static_cast<uint32_t>(f.regions);
      };
      if (!complete()) {
        publish();
        const int64_t t0 = wall_clock64();
        bool stuck = false;
        while (!complete()) {
          if (wall_clock64() - t0 > wait_ticks) {
            stuck = true;
            break;
          }
          __builtin_amdgcn_s_sleep(4);
        }
        if (stuck) {
          if (lane == 0) atomicOr(failed, 1u);
          break;
        }
        fenced_prev = -1;  // a completion newer than the last acquire
      }
      if (seg.prev != fenced_prev) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // see the previous pass's stores
        fenced_prev = seg.prev;
      }
    }
    if (seg.pass > 0) {
      __builtin_amdgcn_wave_barrier();
      if (lane < 16) delta[lane] = move_delta(lane, f.g, BUF ? sizeof(T) : 1u);
      __builtin_amdgcn_wave_barrier();
    }
    const int64_t local = item - seg.item0;
    const int32_t ry = static_cast<int32_t>(local / f.regions_x);
    const int32_t rx = static_cast<int32_t>(local - static_cast<int64_t>(ry) * f.regions_x);
    const int32_t i_base = (ry << (sq_log2 + 3)) + py, j_base = (rx << (sq_log2 + 3)) + px;
    const T* __restrict__ kern = kpool + f.kernel_off;  // uniform: weights via scalar loads
    T* dst = static_cast<T*>(seg.dst);
    if (seg.pass == 0) {
      // one specialised step loop per (uv_mode, recording) pair, each with
      // its own loop over the region (the branches are uniform)
      const bool pol = f.uv_mode == 1, rec = seg.record != 0;
      if (pol && rec) batch_trace_region<T, FMA, BRANCHLESS, 1, 1, BUF>(f, kern, dst, i_base,
                                                                           j_base, sq_log2);
      else if (pol) batch_trace_region<T, FMA, BRANCHLESS, 1, 0, BUF>(f, kern, dst, i_base,
                                                                         j_base, sq_log2);
      else if (rec) batch_trace_region<T, FMA, BRANCHLESS, 0, 1, BUF>(f, kern, dst, i_base,
                                                                         j_base, sq_log2);
      else batch_trace_region<T, FMA, BRANCHLESS, 0, 0, BUF>(f, kern, dst, i_base,
                                                                j_base, sq_log2);
    } else {
      batch_replay_region<T, FMA, BUF>(f, kern, static_cast<const T*>(seg.src), dst, delta,
                                       i_base, j_base, sq_log2);
    }
    if (DEPS) note_done(lo);
  }
  if (DEPS) publish();
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

template <typename T, bool DEPS>
static const void* persistent_fn(int profile, bool buf) {
  const bool fma = profile & 1, bl = profile & 2;
#define RLIC_PICK(BUF)                                                                           \
  (fma && bl ? reinterpret_cast<const void*>(&lic_batch_persistent<T, true, true, BUF, DEPS>)    \
   : fma     ? reinterpret_cast<const void*>(&lic_batch_persistent<T, true, false, BUF, DEPS>)   \
   : bl      ? reinterpret_cast<const void*>(&lic_batch_persistent<T, false, true, BUF, DEPS>)   \
             : reinterpret_cast<const void*>(&lic_batch_persistent<T, false, false, BUF, DEPS>))
  return buf ? RLIC_PICK(true) : RLIC_PICK(false);
#undef RLIC_PICK
}

// Blocks of a persistent grid that fill the device: CUs x resident blocks per CU.
static int persistent_grid(const void* fn) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1024;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0) != hipSuccess ||
      per_cu < 1) {
    (void)hipGetLastError();
    per_cu = 4;
  }
  return cus * per_cu;
}

static size_t align_up(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

// Work-item region side: RLIC_BATCH_REGION (16, 32 or 64), default 32.
// Measured on the 512-field batches: 16 and 32 tie on small fields (sides
// 32..256), 32 is 4% faster than 16 on large ones (sides 128..2048), 64 is
// slower on both.
template <typename T>
static int32_t region_sq_log2(const std::vector<BatchField<T>>&) {
  const char* env = std::getenv("RLIC_BATCH_REGION");
  const int side = (env && *env) ? std::atoi(env) : 32;
  return side == 16 ? 1 : side == 64 ? 3 : 2;
}

template <typename T>
int launch_queue(std::vector<BatchField<T>>& fields, const std::vector<const T*>& kernel_src,
                 std::vector<Segment>& segs, bool deps, bool pack, int profile,
                 hipStream_t stream, std::string& err) {
  const size_t nfields = fields.size();
  if (nfields == 0) return RLIC_OK;
  const int32_t sq_log2 = region_sq_log2(fields);
  const int32_t region = 8 << sq_log2;
  // per-field regions, pack blocks and weight-pool slots
  int64_t pack_blocks = 0, pool_len = 0;
  bool buf = true;
  for (BatchField<T>& f : fields) {
    f.regions_x = (f.g.nx + region - 1) / region;
    f.regions = f.regions_x * ((f.g.ny + region - 1) / region);
    f.pack_block0 = pack_blocks;
    // an empty image has no cells to pack and nothing to copy: its arrays may
    // be zero-size (null pointers), so it gets no pack blocks at all
    if (pack && f.g.ny > 0 && f.g.nx > 0)
      pack_blocks += static_cast<int64_t>((f.g.nx + 2 + 63) / 64) * ((f.g.ny + 2 + 3) / 4);
    f.kernel_off = pool_len + kKernelPad;
    pool_len += f.klen + 2 * kKernelPad;
    if (guarded_cells(f.g.ny, f.g.nx) * static_cast<int64_t>(sizeof(Cell<T>)) > 0xFFFFFFFFll)
      buf = false;
  }
  int64_t items = 0;
  for (Segment& sg : segs) {
    sg.item0 = items;
    items += fields[sg.field].regions;
  }
  // one scratch block: fields | kernel pointers | segments | counters | weight pool
  const size_t off_fields = 0;
  const size_t off_kptrs = align_up(sizeof(BatchField<T>) * nfields);
  const size_t off_segs = align_up(off_kptrs + sizeof(const T*) * nfields);
  const size_t off_done = align_up(off_segs + sizeof(Segment) * std::max<size_t>(segs.size(), 1));
  const size_t n_counters = segs.size() + 4;  // done per segment, queue head (8 B), failure flag
  const size_t off_pool = align_up(off_done + sizeof(uint32_t) * n_counters);
  const size_t bytes = off_pool + sizeof(T) * static_cast<size_t>(pool_len);
  CallScratch scratch;  // descriptors: reused from the per-stream cache
  if (hipError_t e = scratch.get(bytes, stream); e != hipSuccess) {
    err = std::string("allocating queue descriptors: ") + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? RLIC_ERR_OUT_OF_MEMORY : RLIC_ERR_HIP;
  }
  char* base = static_cast<char*>(scratch.ptr());
  T* d_pool = reinterpret_cast<T*>(base + off_pool);
  for (BatchField<T>& f : fields) f.kernel = d_pool + f.kernel_off;
  std::vector<char> host(off_done);
  std::memcpy(host.data() + off_fields, fields.data(), sizeof(BatchField<T>) * nfields);
  std::memcpy(host.data() + off_kptrs, kernel_src.data(), sizeof(const T*) * nfields);
  if (!segs.empty())
    std::memcpy(host.data() + off_segs, segs.data(), sizeof(Segment) * segs.size());
  hipError_t e = hipMemcpyAsync(base, host.data(), off_done, hipMemcpyHostToDevice, stream);
  if (e != hipSuccess) {
    err = std::string("uploading queue descriptors: ") + hipGetErrorString(e);
    return RLIC_ERR_HIP;
  }
  auto* d_fields = reinterpret_cast<const BatchField<T>*>(base + off_fields);
  auto* d_kptrs = reinterpret_cast<const T* const*>(base + off_kptrs);
  auto* d_segs = reinterpret_cast<const Segment*>(base + off_segs);
  auto* d_done = reinterpret_cast<uint32_t*>(base + off_done);
  auto* d_queue = reinterpret_cast<unsigned long long*>(d_done + segs.size() + (segs.size() & 1));
  auto* d_failed = reinterpret_cast<uint32_t*>(d_queue + 1);

  hipLaunchKernelGGL(lic_zero_u32, dim3(static_cast<unsigned>(std::min<size_t>(
                                       (n_counters + 255) / 256, 1024))),
                     dim3(256), 0, stream, d_done, static_cast<int64_t>(n_counters));
  hipLaunchKernelGGL(lic_pad_batch<T>, dim3(static_cast<unsigned>(nfields)), dim3(256), 0, stream,
                     d_fields, d_kptrs);
  if (pack && pack_blocks > 0) {
    hipLaunchKernelGGL(lic_pack_batch<T>, dim3(static_cast<unsigned>(pack_blocks)), dim3(256), 0,
                       stream, d_fields, static_cast<int32_t>(nfields));
  }
  if ((e = hipGetLastError()) != hipSuccess) {
    err = std::string("queue set-up kernels: ") + hipGetErrorString(e);
    return RLIC_ERR_HIP;
  }
  if (items > 0) {
    // four waves (workers) per block; with deps, a wait is declared stuck
    // after RLIC_BATCH_WAIT_MS milliseconds (default 2000) of the device's
    // wall clock
    const void* fn = deps ? persistent_fn<T, true>(profile, buf) : persistent_fn<T, false>(profile, buf);
    const int grid =
        static_cast<int>(std::min<int64_t>((items + 3) / 4, persistent_grid(fn)));
    int dev = 0, khz = 100000;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess ||
        khz <= 0) {
      (void)hipGetLastError();
      khz = 100000;
    }
    const char* env = std::getenv("RLIC_BATCH_WAIT_MS");
    const int64_t wait_ticks = (env ? std::max<int64_t>(0, std::atoll(env)) : 2000) * khz;
    const int32_t nsegs = static_cast<int32_t>(segs.size());
    int32_t sq = sq_log2;
    void* args[] = {&d_fields, &d_segs, &d_pool, const_cast<int32_t*>(&nsegs), &items,
                    &d_queue, &d_done, &d_failed, const_cast<int64_t*>(&wait_ticks), &sq};
    e = hipLaunchKernel(fn, dim3(grid), dim3(256), args, 0, stream);
    if (e != hipSuccess) {
      err = std::string("queue kernel: ") + hipGetErrorString(e);
      return RLIC_ERR_HIP;
    }
  }
  uint32_t failed = 0;
  e = hipMemcpyAsync(&failed, d_failed, sizeof(failed), hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) {
    err = std::string("queue completion: ") + hipGetErrorString(e);
    return RLIC_ERR_HIP;
  }
  if (failed) {
    err = "a work item waited too long for the previous pass of its field";
    return RLIC_ERR_HIP;
  }
  return RLIC_OK;
}

// One field per grid column (blockIdx.x), gridDim.y blocks down the rows
// of its texture (strided, as given), 256 lanes along each row; block 0
// also checks the kernel weights.  flags[2k] = 1: field k has a texture
// element below 0 (a NaN element passes, as in the reference: `texture <
// 0` is false for NaN); flags[2k + 1] = 1: a non-finite weight.  Every
// writer of a word stores the same value.  HBM-bound: one read of every
// texture element and weight.
template <typename T>
__device__ __forceinline__ void check_field(const CheckField<T>& f, uint32_t* flags, uint32_t k) {
  // no early exit: four independent loads in flight per lane
  bool neg = false;
  for (int64_t i = blockIdx.y; i < f.ny; i += gridDim.y) {
    const T* row = f.texture + i * f.tex_sy;
    int64_t j = threadIdx.x;
    for (; j + 768 < f.nx; j += 1024) {
      const T a = row[j * f.tex_sx], b = row[(j + 256) * f.tex_sx];
      const T c = row[(j + 512) * f.tex_sx], d = row[(j + 768) * f.tex_sx];
      neg |= (a < T(0)) | (b < T(0)) | (c < T(0)) | (d < T(0));
    }
    for (; j < f.nx; j += 256) neg |= row[j * f.tex_sx] < T(0);
  }
  if (neg) flags[2 * k] = 1u;
  if (blockIdx.y == 0) {
    bool bad = false;
    for (int64_t t = threadIdx.x; t < f.kernel_len; t += 256) bad |= !__builtin_isfinite(f.kernel[t]);
    if (bad) flags[2 * k + 1] = 1u;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void lic_check_values(const CheckField<T>* __restrict__ table,
                                                        uint32_t* flags) {
  check_field(table[blockIdx.x], flags, blockIdx.x);
}

// Up to kCheckInline fields passed in the kernel arguments (no table copy).
template <typename T>
__global__ __launch_bounds__(256) void lic_check_values_inline(const CheckPack<T> pack,
                                                               uint32_t* flags) {
  check_field(pack.f[blockIdx.x], flags, blockIdx.x);
}

template <typename T>
hipError_t launch_check_values(const CheckField<T>* table, const CheckPack<T>* pack,
                               int32_t nfields, int64_t max_pixels, uint32_t* flags,
                               hipStream_t stream) {
  if (nfields <= 0) return hipSuccess;
  // about 4 elements per lane of the largest field (a small field is
  // latency-bound: few loads per lane), at most 4096 blocks per field
  // (blocks of smaller fields past their last row exit at once)
  const int64_t per_block = 256 * 4;
  const auto blocks = static_cast<uint32_t>(
      std::min<int64_t>(4096, std::max<int64_t>(1, (max_pixels + per_block - 1) / per_block)));
  const dim3 grid(static_cast<uint32_t>(nfields), blocks);
  if (pack) {
    if (nfields > kCheckInline) return hipErrorInvalidValue;
    hipLaunchKernelGGL((lic_check_values_inline<T>), grid, dim3(256), 0, stream, *pack, flags);
  } else {
    hipLaunchKernelGGL((lic_check_values<T>), grid, dim3(256), 0, stream, table, flags);
  }
  return hipGetLastError();
}

template <typename T, typename FieldSpec>
int batch_run(const FieldSpec* specs, int64_t nfields, int profile, hipStream_t stream,
              BatchScratch& scratch, std::string& err) {
  // per-field working memory: cells, ping-pong partner, move record
  std::vector<size_t> off_cells(nfields), off_tmp(nfields), off_codes(nfields);
  size_t bytes = 0;
  for (int64_t k = 0; k < nfields; ++k) {
    const FieldSpec& s = specs[k];
    const size_t n = static_cast<size_t>(s.ny * s.nx);
    off_cells[k] = bytes;
    if (s.iterations > 0)
      bytes = align_up(bytes + sizeof(Cell<T>) * static_cast<size_t>(guarded_cells(s.ny, s.nx)));
    off_tmp[k] = bytes;
    if (s.iterations > 1) bytes = align_up(bytes + sizeof(T) * n);
    off_codes[k] = bytes;
    if (s.iterations > 1)
      bytes = align_up(bytes + sizeof(uint4) * code_chunks(static_cast<int32_t>(s.kernel_len)) * n);
  }
  if (hipError_t e = scratch.reserve(bytes, stream); e != hipSuccess) {
    err = std::string("allocating batch working memory: ") + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? RLIC_ERR_OUT_OF_MEMORY : RLIC_ERR_HIP;
  }
  char* base = static_cast<char*>(scratch.ptr);
  std::vector<BatchField<T>> fields(nfields);
  std::vector<const T*> ksrc(nfields);
  std::vector<T*> tmp(nfields);
  int64_t max_it = 0;
  for (int64_t k = 0; k < nfields; ++k) {
    const FieldSpec& s = specs[k];
    BatchField<T>& f = fields[k];
    std::memset(&f, 0, sizeof(f));
    f.tex = s.texture; f.tsy = s.tex_sy; f.tsx = s.tex_sx;
    f.u = s.u; f.usy = s.u_sy; f.usx = s.u_sx;
    f.v = s.v; f.vsy = s.v_sy; f.vsx = s.v_sx;
    f.out = s.out;
    f.cells = s.iterations > 0 ? reinterpret_cast<Cell<T>*>(base + off_cells[k]) : nullptr;
    f.codes = s.iterations > 1 ? reinterpret_cast<uint4*>(base + off_codes[k]) : nullptr;
    f.g = FieldGeom{static_cast<int32_t>(s.ny), static_cast<int32_t>(s.nx), s.x_left, s.x_right,
                    s.y_left, s.y_right};
    f.klen = static_cast<int32_t>(s.kernel_len);
    f.uv_mode = s.uv_mode;
    f.iterations = static_cast<int32_t>(s.iterations);
    ksrc[k] = s.kernel;
    tmp[k] = s.iterations > 1 ? reinterpret_cast<T*>(base + off_tmp[k]) : nullptr;
    max_it = std::max<int64_t>(max_it, s.iterations);
  }
  // pass-major segments; pass p of a field with n passes writes `out` when
  // n-1-p is even and the ping-pong partner otherwise, so the last pass
  // lands in `out`
  std::vector<Segment> segs;
  std::vector<int32_t> last_seg(nfields, -1);
  for (int64_t pass = 0; pass < max_it; ++pass) {
    for (int64_t k = 0; k < nfields; ++k) {
      const FieldSpec& s = specs[k];
      if (s.iterations <= pass || s.ny * s.nx == 0) continue;
      auto buffer = [&](int64_t p) -> T* { return ((s.iterations - 1 - p) % 2 == 0) ? s.out : tmp[k]; };
      Segment sg{};
      sg.field = static_cast<int32_t>(k);
      sg.pass = static_cast<int32_t>(pass);
      sg.prev = last_seg[k];
      sg.record = (pass == 0 && s.iterations > 1) ? 1 : 0;
      sg.src = pass > 0 ? buffer(pass - 1) : nullptr;
      sg.dst = buffer(pass);
      last_seg[k] = static_cast<int32_t>(segs.size());
      segs.push_back(sg);
    }
  }
  return launch_queue<T>(fields, ksrc, segs, /*deps=*/true, /*pack=*/true, profile, stream, err);
}

template int launch_queue<float>(std::vector<BatchField<float>>&, const std::vector<const float*>&,
                                 std::vector<Segment>&, bool, bool, int, hipStream_t,
                                 std::string&);
template int launch_queue<double>(std::vector<BatchField<double>>&,
                                  const std::vector<const double*>&, std::vector<Segment>&, bool,
                                  bool, int, hipStream_t, std::string&);
template hipError_t launch_check_values<float>(const CheckField<float>*, const CheckPack<float>*,
                                               int32_t, int64_t, uint32_t*, hipStream_t);
template hipError_t launch_check_values<double>(const CheckField<double>*,
                                                const CheckPack<double>*, int32_t, int64_t,
                                                uint32_t*, hipStream_t);
template int batch_run<float, rlic_field_f32>(const rlic_field_f32*, int64_t, int, hipStream_t,
                                              BatchScratch&, std::string&);
template int batch_run<double, rlic_field_f64>(const rlic_field_f64*, int64_t, int, hipStream_t,
                                               BatchScratch&, std::string&);

}  // namespace rlic
