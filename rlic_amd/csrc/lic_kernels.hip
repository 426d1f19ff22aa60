// HIP kernels of the line integral convolution hot path (gfx950).
//
//   lic_pack    strided (texture, u, v) -> guarded Cell array (one pass, HBM-bound)
//   lic_trace   one lane per output pixel: centre tap, forward streamline,
//               backward streamline (src/lib.rs:364-406, 304-362); optionally
//               records the cell-to-cell move of every step
//   lic_replay  later iterations: the vector field does not change between
//               iterations (src/lib.rs:416-420, 432-440), so every iteration
//               visits the same cells; replay walks the recorded moves and
//               only gathers the new texture, in the same summation order
//   lic_copy2d  strided 2-D copy (iterations == 0)
//   lic_repack_texture  texture slot refresh when passes are re-traced
#include <algorithm>
#include <type_traits>

#include "lic_device.h"
#include "lic_launch.h"
#include "lic_steps.h"

namespace rlic {

// The guarded cell array of a field (guarded_cells): the image's cells at
// (i + 1, j + 1) and a ring of guard cells around them.  A step off the
// image lands on a guard cell, which holds the values of the cell the
// boundary rule sends that step to (guard_source) and, in its pad slot, the
// jump in cells from the guard cell to that cell (0 in image cells).
// Corner cells are never reached and stay zero.  With `kpad`, the same
// launch also writes the padded weight copy [pad zeros][kernel][pad zeros]
// (every thread of the grid takes a share), which saves a small field's
// call one launch.
template <typename T>
__global__ __launch_bounds__(256) void lic_pack(const T* __restrict__ tex, int64_t tsy,
                                                int64_t tsx, const T* __restrict__ u,
                                                int64_t usy, int64_t usx,
                                                const T* __restrict__ v, int64_t vsy,
                                                int64_t vsx, FieldGeom g,
                                                Cell<T>* __restrict__ cells,
                                                const T* __restrict__ kernel, int32_t klen,
                                                T* __restrict__ kpad) {
  if (kpad) {
    const int64_t threads = static_cast<int64_t>(gridDim.x) * gridDim.y * 256;
    const int64_t total = int64_t{klen} + 2 * kKernelPad;
    for (int64_t t = (static_cast<int64_t>(blockIdx.y) * gridDim.x + blockIdx.x) * 256 +
                     threadIdx.x;
         t < total; t += threads) {
      const int64_t k = t - kKernelPad;
      kpad[t] = (k >= 0 && k < klen) ? kernel[k] : T(0);
    }
  }
  const int32_t gj = blockIdx.x * 64 + (threadIdx.x & 63);
  const int32_t gi = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (gi >= g.ny + 2 || gj >= g.nx + 2) return;
  cells[static_cast<int64_t>(gi) * (g.nx + 2) + gj] =
      guarded_cell(tex, tsy, tsx, u, usy, usx, v, vsy, vsx, g, gi, gj);
}

// The texture slot of every cell of the guarded array from the C-order
// image x (guard cells from their boundary-rule source).
template <typename T>
__global__ __launch_bounds__(256) void lic_repack_texture(const T* __restrict__ x, FieldGeom g,
                                                          Cell<T>* __restrict__ cells) {
  const int32_t gj = blockIdx.x * 64 + (threadIdx.x & 63);
  const int32_t gi = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (gi >= g.ny + 2 || gj >= g.nx + 2) return;
  int32_t i, j;
  if (guard_source(gi, gj, g.ny, g.nx, g.x_left, g.x_right, g.y_left, g.y_right, i, j))
    cells[static_cast<int64_t>(gi) * (g.nx + 2) + gj].x = x[static_cast<int64_t>(i) * g.nx + j];
}

template <typename T>
__global__ __launch_bounds__(256) void lic_copy2d(const T* __restrict__ src, int64_t sy,
                                                  int64_t sx, int32_t ny, int32_t nx,
                                                  T* __restrict__ dst) {
  const int32_t j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int32_t i = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= ny || j >= nx) return;
  dst[static_cast<int64_t>(i) * nx + j] = src[i * sy + j * sx];
}

// The cells are the field's guarded array (lic_pack); a position is a
// byte offset into it with BUF (buffer loads, arrays below 4 GiB), else a
// cell index.
template <typename T, bool FMA, bool BRANCHLESS, bool POL, bool REC, bool BUF>
__global__ __launch_bounds__(256) void lic_trace(const Cell<T>* __restrict__ cells,
                                                 const T* __restrict__ kernel, int32_t klen,
                                                 FieldGeom g, TileMap m, T* __restrict__ out,
                                                 uint4* __restrict__ codes) {
  int32_t ty, tx, py, px;
  tile_of_block(blockIdx.x, m, ty, tx);
  pixel_of_thread(threadIdx.x, m, py, px);
  const int32_t i0 = ty * tile_h(m) + py, j0 = tx * tile_w(m) + px;
  if (i0 >= g.ny || j0 >= g.nx) return;
  const CellGather<T, BUF> cell(cells, guarded_cells(g.ny, g.nx));
  const StepGeom sg = make_step_geom(g, cell.unit);
  const int32_t kmid = klen / 2;
  const int32_t n_fwd = klen - 1 - kmid, n_bwd = kmid;
  const int64_t np = static_cast<int64_t>(g.ny) * g.nx;  // code slot = pixel index
  const int64_t p = static_cast<int64_t>(i0) * g.nx + j0;
  const uint32_t row0 =
      static_cast<uint32_t>(i0 + 1) * sg.grow + static_cast<uint32_t>(j0 + 1) * cell.unit;
  const CellVal<T> c0 = cell(row0);
  T acc = macc<FMA>(kernel[kmid], c0.x, T(0));  // src/lib.rs:375-383
  acc = trace_direction<T, FMA, BRANCHLESS, POL, true, REC, BUF>(
      acc, c0.u, c0.v, row0, cell, kernel, kmid, n_fwd, sg, codes, np, p, 0);
  acc = trace_direction<T, FMA, BRANCHLESS, POL, false, REC, BUF>(
      acc, c0.u, c0.v, row0, cell, kernel, kmid, n_bwd, sg, codes, np, p,
      chunks_of(n_fwd));
  out[static_cast<int64_t>(i0) * g.nx + j0] = acc;
}

// Paired trace for fields too small to fill the GPU: the two directions of
// a pixel's line run on two lanes of one wave at once (lanes 0..31 forward,
// 32..63 backward, of the same 32 pixels), so a lane takes half the
// dependent steps of lic_trace and a field has twice the waves.  The
// forward lane accumulates as lic_trace does; the backward lane keeps the
// texture values of its live steps in LDS, which the forward lane then adds
// in the reference's order (centre, forward steps, backward steps:
// src/lib.rs:375-383), so the sum is lic_trace's bit for bit.  With REC,
// each lane records its own direction.  Needs klen odd (both directions
// kmid steps) and LDS for 32 x (kmid + 1) values (trace_pair_fits).  Blocks
// of one wave, eight per tile of the TileMap (lic_trace's four waves of 64
// pixels as eight of 32): one-wave blocks spread a small field's waves over
// the CUs (measured faster than 512-thread blocks, one per tile, from 32^2
// to 362^2 pixels).
template <typename T, bool FMA, bool BRANCHLESS, bool POL, bool REC>
__global__ __launch_bounds__(64) void lic_trace_pair(
    const Cell<T>* __restrict__ cells, const T* __restrict__ kernel, int32_t klen, FieldGeom g,
    TileMap m, T* __restrict__ out, uint4* __restrict__ codes) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int32_t lane = threadIdx.x;
  const bool fwd = lane < 32;
  const int32_t k = lane & 31;
  // a block is one wave: eight blocks take the 256 pixels of a tile
  int32_t ty, tx, py, px;
  tile_of_block(static_cast<int32_t>(blockIdx.x) >> 3, m, ty, tx);
  pixel_of_thread((static_cast<int32_t>(blockIdx.x) & 7) * 32 + k, m, py, px);
  const int32_t i0 = ty * tile_h(m) + py, j0 = tx * tile_w(m) + px;
  const bool inside = i0 < g.ny && j0 < g.nx;
  const int32_t kmid = klen / 2;  // steps of either direction
  // the backward value of step s of pixel k at vals[32 s] (rows 0..kmid - 1;
  // row kmid takes the forward lanes' stores)
  typedef T __attribute__((address_space(3))) lds_t;
  lds_t* vals = (lds_t*)smem + k;
  const int64_t p = static_cast<int64_t>(i0) * g.nx + j0;
  T acc = T(0);
  int32_t live_steps = 0;  // backward lanes: steps before the line ended
  if (inside) {
    const CellGather<T, true> cell(cells, guarded_cells(g.ny, g.nx));
    const StepGeom sg = make_step_geom(g, cell.unit);
    const int64_t np = static_cast<int64_t>(g.ny) * g.nx;
    const uint32_t row0 =
        static_cast<uint32_t>(i0 + 1) * sg.grow + static_cast<uint32_t>(j0 + 1) * cell.unit;
    const CellVal<T> c0 = cell(row0);
    acc = macc<FMA>(kernel[kmid], c0.x, T(0));  // src/lib.rs:375-383
    // every lane stores (no branch in the step): forward lanes into the
    // spare row kmid
    auto keep = [&](int32_t s, T x, bool live) {
      vals[32 * (fwd ? kmid : s)] = x;
      live_steps = live ? s + 1 : live_steps;
    };
    acc = trace_direction<T, FMA, BRANCHLESS, POL, true, REC, true, true>(
        acc, c0.u, c0.v, row0, cell, kernel, kmid, kmid, sg, codes, np, p,
        fwd ? 0 : chunks_of(kmid), false, false, fwd, keep);
  }
  __syncthreads();
  const int32_t nb = __shfl(live_steps, k + 32);  // the backward partner's
  if (fwd && inside) {
    for (int32_t s = 0; s < kmid; ++s) {
      const T next_acc = macc<FMA>(kernel[kmid - 1 - s], vals[32 * s], acc);
      acc = s < nb ? next_acc : acc;
    }
    out[p] = acc;
  }
}

// Lane work of a tracing pass, read back from its move record.  The trace
// kernel runs a wave in groups of 8 steps until every lane's line has ended
// (or the steps run out); a lane whose line has ended keeps executing the
// group with its position frozen.  Per wave and direction this kernel finds
// each lane's first stop code (the step where its line ended), and adds
// the lanes' live steps to stats[0] and the steps the wave executed, times
// the lanes that hold a pixel, to stats[1].  Same block and lane mapping as
// lic_trace.
__global__ __launch_bounds__(256) void lic_record_stats(const uint4* __restrict__ codes,
                                                        int32_t klen, FieldGeom g, TileMap m,
                                                        unsigned long long* __restrict__ stats) {
  int32_t ty, tx, py, px;
  tile_of_block(blockIdx.x, m, ty, tx);
  pixel_of_thread(threadIdx.x, m, py, px);
  const int32_t i0 = ty * tile_h(m) + py, j0 = tx * tile_w(m) + px;
  const bool has_pixel = i0 < g.ny && j0 < g.nx;
  const int64_t np = static_cast<int64_t>(g.ny) * g.nx;
  const int64_t p = has_pixel ? static_cast<int64_t>(i0) * g.nx + j0 : 0;
  const int32_t kmid = klen / 2;
  const int32_t lanes = __popcll(__builtin_amdgcn_ballot_w64(has_pixel));
  unsigned long long useful = 0, executed = 0;
  for (int dir = 0; dir < 2; ++dir) {
    const int32_t nsteps = dir == 0 ? klen - 1 - kmid : kmid;
    const int32_t c0 = dir == 0 ? 0 : chunks_of(klen - 1 - kmid);
    int32_t end = nsteps;  // first step whose code is the stop code
    if (has_pixel) {
      for (int32_t ch = 0; ch < chunks_of(nsteps) && end == nsteps; ++ch) {
        const uint4 c = codes[(c0 + ch) * np + p];
        const uint32_t words[4] = {c.x, c.y, c.z, c.w};
        for (int wi = 0; wi < 4 && end == nsteps; ++wi) {
          for (int q = 0; q < 8; ++q) {
            const int32_t s = ch * kChunkSteps + 8 * wi + q;
            if (s < nsteps && move_code(words[wi], q) == kMoveStop) {
              end = s;
              break;
            }
          }
        }
      }
    }
    if (has_pixel) useful += static_cast<unsigned long long>(end);
    // the wave's last group is the one where its last line ended
    int32_t last_group = (has_pixel && end == nsteps) ? INT32_MAX : (has_pixel ? end / 8 : -1);
    for (int off = 32; off > 0; off >>= 1) last_group = max(last_group, __shfl_xor(last_group, off));
    const int64_t wave_steps =
        last_group < 0 ? 0 : min<int64_t>(nsteps, 8 * (static_cast<int64_t>(last_group) + 1));
    executed += static_cast<unsigned long long>(wave_steps) * lanes;
  }
  for (int off = 32; off > 0; off >>= 1) useful += __shfl_xor(useful, off);
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&stats[0], useful);
    atomicAdd(&stats[1], executed);
  }
}

template <typename T, bool FMA, bool BUF>
__global__ __launch_bounds__(256) void lic_replay(const T* __restrict__ xin,
                                                  const T* __restrict__ kernel, int32_t klen,
                                                  FieldGeom g, TileMap m, T* __restrict__ out,
                                                  const uint4* __restrict__ codes) {
  const ValueGather<T, BUF> x(xin, static_cast<int64_t>(g.ny) * g.nx);
  __shared__ uint32_t delta[16];
  if (threadIdx.x < 16) delta[threadIdx.x] = move_delta(threadIdx.x, g, x.unit);
  __syncthreads();
  int32_t ty, tx, py, px;
  tile_of_block(blockIdx.x, m, ty, tx);
  pixel_of_thread(threadIdx.x, m, py, px);
  const int32_t i0 = ty * tile_h(m) + py, j0 = tx * tile_w(m) + px;
  if (i0 >= g.ny || j0 >= g.nx) return;
  const int32_t kmid = klen / 2;
  const int32_t n_fwd = klen - 1 - kmid, n_bwd = kmid;
  const int64_t np = static_cast<int64_t>(g.ny) * g.nx;  // code slot = pixel index
  const int64_t idx0 = static_cast<int64_t>(i0) * g.nx + j0;
  const int64_t p = idx0;
  const uint32_t pos0 = static_cast<uint32_t>(idx0) * x.unit;
  T acc = macc<FMA>(kernel[kmid], x(pos0), T(0));
  acc = replay_direction<T, FMA, true, BUF>(acc, pos0, x, kernel, kmid, n_fwd, delta, codes, np,
                                            p, 0);
  acc = replay_direction<T, FMA, false, BUF>(acc, pos0, x, kernel, kmid, n_bwd, delta, codes,
                                             np, p, chunks_of(n_fwd));
  out[idx0] = acc;
}

// ---------------------------------------------------------------------------
// LDS-staged replay.  A block of 1024 threads replays a TH x TW pixel tile
// (TH, TW in {32, 64}), E = TH * TW / 1024 pixels per lane.  Every step
// moves one cell, so all lines of the tile stay inside the tile grown by
// kmid cells on each side (in unwrapped coordinates: a periodic edge is
// just crossed).  The block stages that box of the texture into LDS once
// (wrapped modulo the image size), then replays every line from LDS, where
// a move is +-1 or +-stride whatever the boundary did (a clamp was recorded
// as "stay").
//
// LDS layout: the row stride is 8 (mod 32) elements, which puts the 64
// lanes of a wave's 8 x 8 pixel square in 64 distinct 4-byte banks
// (32-bank and 64-bank LDS alike), so an unsheared square reads without
// bank conflicts.  Moves are decoded two at a time through a 256-entry
// table indexed by a code byte (two consecutive steps, the earlier in the
// high nibble), holding the two byte-offset deltas as 32-bit integers (one
// add per step).  The pair table is the block's only
// static LDS, at address 0 (a table address is the code byte times 8); the
// dynamic LDS after it (a multiple of 64 banks further) holds the box.
// ---------------------------------------------------------------------------
constexpr int32_t kLdsThreads = 1024;

// Box of a tile grown by `halo` cells on each side (halo = kmid: every
// step of every line; smaller: the first `halo` steps, see the hybrid
// replay below).
__host__ __device__ inline int32_t lds_box_h(int32_t th, int32_t halo) { return th + 2 * halo; }
__host__ __device__ inline int32_t lds_box_w(int32_t tw, int32_t halo) { return tw + 2 * halo; }
__host__ __device__ inline int32_t lds_box_stride(int32_t tw, int32_t halo) {
  const int32_t w = lds_box_w(tw, halo);
  return w + ((8 - w) % 32 + 32) % 32;  // == 8 (mod 32)
}

constexpr size_t kLdsPairTableBytes = 256 * sizeof(int2);  // static LDS, per table

// Dynamic LDS of one block: the box.
inline size_t lds_replay_dynamic_bytes(int32_t th, int32_t tw, int32_t halo, size_t elem) {
  return static_cast<size_t>(lds_box_h(th, halo)) * lds_box_stride(tw, halo) * elem;
}

// All LDS of one block (`tables` pair tables: 1, or 2 for the hybrid replay).
inline size_t lds_replay_bytes(int32_t th, int32_t tw, int32_t halo, size_t elem,
                               int32_t tables = 1) {
  return tables * kLdsPairTableBytes + lds_replay_dynamic_bytes(th, tw, halo, elem);
}

// Hybrid replay: a 64 x 64 tile whose box holds the first `halo` steps of
// each direction (32 or 64, a multiple of the 32-step chunk); later steps
// gather from the image through buffer loads.  The largest halo whose box
// fits (hybrid_max_halo): 64 in float (192 x 200 x 4 B), 32 in double
// (128 x 136 x 8 B).

__device__ __forceinline__ int32_t wrap_index(int32_t c, int32_t n) {
  c %= n;
  return c < 0 ? c + n : c;
}

// Byte-offset change in the LDS box for one move code: in the box's
// unwrapped coordinates a periodic edge step is an ordinary step.
__device__ __forceinline__ int32_t lds_move(uint32_t c, const FieldGeom& g, int32_t row_bytes,
                                            int32_t elem) {
  if ((c & kMoveMoved) == 0) return 0;  // stay, stop (and the unused codes)
  const bool y = (c & kMoveAxisY) != 0, neg = (c & kMoveNeg) != 0;
  const bool per = y ? (neg ? g.y_left : g.y_right) : (neg ? g.x_left : g.x_right);
  if ((c & kMoveEdge) != 0 && !per) return 0;  // a closed edge keeps the cell
  const int32_t d = y ? row_bytes : elem;
  return neg ? -d : d;
}

// A line's position while replaying: an LDS byte pointer into the staged
// box (LdsWalk) or a byte offset into the image, read with buffer loads
// (ImageWalk, the hybrid replay's steps past the box).  Either way each
// step is one add of a pair-table delta.
typedef const unsigned char __attribute__((address_space(3))) lds_byte;

template <typename T>
__device__ __forceinline__ T lds_value(lds_byte* at) {
  typedef const T __attribute__((address_space(3))) lds_t;
  return *reinterpret_cast<lds_t*>(at);
}

// kTableFirst: the replay loop reads the pair-table entries of all of a
// lane's lines before their gathers (LDS gathers are short; the early table
// reads hide their latency), or each line's just before its own gathers
// (image gathers are long; fewer registers live).
template <typename T>
struct LdsWalk {
  typedef lds_byte* Pos;
  static constexpr bool kTableFirst = true;
  __device__ __forceinline__ T operator()(Pos at) const { return lds_value<T>(at); }
};

template <typename T>
struct ImageWalk {
  typedef uint32_t Pos;
  static constexpr bool kTableFirst = false;
  __amdgpu_buffer_rsrc_t rsrc;  // the image, n elements with n * sizeof(T) < 2^32
  // base and size pass through readfirstlane: the descriptor must be
  // wave-uniform (scalar registers), which the compiler does not prove here
  __device__ __forceinline__ ImageWalk(const T* base, int64_t n) {
    const uint64_t b = reinterpret_cast<uint64_t>(base);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(b));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(b >> 32));
    const uint32_t bytes =
        __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(n * static_cast<int64_t>(sizeof(T))));
    rsrc = make_rsrc(reinterpret_cast<const void*>((uint64_t{hi} << 32) | lo), bytes);
  }
  __device__ __forceinline__ T operator()(Pos at) const {
    return buf_load(rsrc, at, static_cast<T*>(nullptr));
  }
};

// The 8 values one line reads over a group of 8 moves (`word`, one 4-bit
// move code per step), advancing its position `at`.
template <typename T, typename Walk>
__device__ __forceinline__ void walk_group(const Walk& walk, T (&xv)[8],
                                           typename Walk::Pos& at, uint32_t word,
                                           const int2* pair_delta) {
#pragma unroll
  for (int h = 0; h < 4; ++h) {
    const int2 d = pair_delta[(word >> (24 - 8 * h)) & 255u];
    at += d.x;
    xv[2 * h] = walk(at);
    at += d.y;
    xv[2 * h + 1] = walk(at);
  }
}

// Replays chunks [ch_begin, min(ch_end, chunks of nsteps)) of one direction
// of E lines per lane; `at` and `alive` carry the lines' state in and out.
template <typename T, bool FMA, bool FORWARD, int E, typename Walk>
__device__ __forceinline__ void replay_lines(
    const Walk& walk, T (&acc)[E], typename Walk::Pos (&at)[E], bool (&alive)[E],
    const T* __restrict__ kernel, int32_t kmid, int32_t nsteps, const int2* pair_delta,
    const uint4* __restrict__ codes, int64_t np, const int64_t (&p)[E], int32_t c0,
    int32_t ch_begin, int32_t ch_end) {
  const int32_t nchunks = min(chunks_of(nsteps), ch_end);
  uint4 next[E];
#pragma unroll
  for (int e = 0; e < E; ++e)
    next[e] = (ch_begin < nchunks && alive[e]) ? codes[(c0 + ch_begin) * np + p[e]]
                                               : make_uint4(0u, 0u, 0u, 0u);
  bool any_line = false;
#pragma unroll
  for (int e = 0; e < E; ++e) any_line = any_line || alive[e];
  if (__builtin_amdgcn_ballot_w64(any_line) == 0) return;
  for (int32_t ch = ch_begin; ch < nchunks; ++ch) {
    uint32_t words[E][4];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      words[e][0] = next[e].x;
      words[e][1] = next[e].y;
      words[e][2] = next[e].z;
      words[e][3] = next[e].w;
      if (ch + 1 < nchunks && alive[e]) next[e] = codes[(c0 + ch + 1) * np + p[e]];
    }
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int32_t s0 = ch * kChunkSteps + 8 * g;
      if (s0 >= nsteps) break;  // uniform
      T kw[8];
#pragma unroll
      for (int q = 0; q < 8; ++q)
        kw[q] = kernel[FORWARD ? kmid + 1 + s0 + q : kmid - 1 - s0 - q];
      const int32_t lim = nsteps - s0;
      uint32_t word[E];
      bool full = lim >= 8;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        word[e] = alive[e] ? words[e][g] : 0u;  // 0 = stay
        full = full && alive[e] && !group_has_stop(word[e]);
      }
      if (__builtin_amdgcn_ballot_w64(!full) == 0) {
        // every lane accumulates all 8 steps of every pixel (wave-uniform):
        // no per-step masks, and every line stays alive.  The pair-table
        // reads go first (see kTableFirst), and line e + 1's gathers are
        // issued before line e accumulates, so gathers overlap.
        int2 d[E][4];
        if constexpr (Walk::kTableFirst) {
#pragma unroll
          for (int e = 0; e < E; ++e)
#pragma unroll
            for (int h = 0; h < 4; ++h) d[e][h] = pair_delta[(word[e] >> (24 - 8 * h)) & 255u];
        }
        T xv[2][8];
#pragma unroll
        for (int e = 0; e <= E; ++e) {
          if (e < E) {
            if constexpr (!Walk::kTableFirst) {
#pragma unroll
              for (int h = 0; h < 4; ++h) d[e][h] = pair_delta[(word[e] >> (24 - 8 * h)) & 255u];
            }
#pragma unroll
            for (int h = 0; h < 4; ++h) {
              at[e] += d[e][h].x;
              xv[e & 1][2 * h] = walk(at[e]);
              at[e] += d[e][h].y;
              xv[e & 1][2 * h + 1] = walk(at[e]);
            }
          }
          if (e > 0) {
#pragma unroll
            for (int q = 0; q < 8; ++q)
              acc[e - 1] = macc<FMA>(kw[q], xv[(e - 1) & 1][q], acc[e - 1]);
          }
        }
      } else {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int32_t valid = alive[e] ? steps_before_stop(word[e]) : 0;
          T xv[8];
          walk_group<T>(walk, xv, at[e], word[e], pair_delta);
          // branch-free over the group (see replay_direction)
          const int32_t use = min(valid, lim);
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const T next_acc = macc<FMA>(kw[q], xv[q], acc[e]);
            acc[e] = q < use ? next_acc : acc[e];
          }
          alive[e] = valid == 8;
        }
      }
    }
    bool any = false;
#pragma unroll
    for (int e = 0; e < E; ++e) any = any || alive[e];
    if (__builtin_amdgcn_ballot_w64(any) == 0) break;
  }
}

template <typename T, bool FMA, int TH, int TW, bool HYBRID>
__global__ __launch_bounds__(kLdsThreads) void lic_replay_lds(const T* __restrict__ x,
                                                              const T* __restrict__ kernel,
                                                              int32_t klen, FieldGeom g,
                                                              TileMap m, T* __restrict__ out,
                                                              const uint4* __restrict__ codes,
                                                              int32_t box_halo) {
  constexpr int E = TH * TW / kLdsThreads;  // pixels per lane: 1, 2 or 4
  constexpr int32_t elem = static_cast<int32_t>(sizeof(T));
  extern __shared__ __align__(16) unsigned char smem[];
  const int32_t kmid = klen / 2;
  // the box holds every step (halo kmid), or with HYBRID the first
  // box_halo steps of each direction (a multiple of 32 below kmid)
  const int32_t halo = HYBRID ? box_halo : kmid;
  const int32_t bh = lds_box_h(TH, halo), bw = lds_box_w(TW, halo);
  const int32_t stride = lds_box_stride(TW, halo);
  __shared__ int2 pair_delta[256];
  __shared__ int2 pair_image[HYBRID ? 256 : 1];  // image-offset deltas (wraps jump)
  unsigned char* box = smem;

  int32_t ty, tx;
  tile_of_block(blockIdx.x, m, ty, tx);
  const int32_t i_box0 = ty * TH - halo, j_box0 = tx * TW - halo;
  const int32_t tid = threadIdx.x;
  const int32_t row_bytes = stride * elem;
  if (tid < 256) {
    // a code byte holds two steps, the earlier in its high nibble
    const int32_t lo = lds_move(tid >> 4, g, row_bytes, elem);
    const int32_t hi = lds_move(tid & 15, g, row_bytes, elem);
    pair_delta[tid] = make_int2(lo, hi);
    if constexpr (HYBRID) {
      pair_image[tid] = make_int2(static_cast<int32_t>(move_delta(tid >> 4, g, elem)),
                                  static_cast<int32_t>(move_delta(tid & 15, g, elem)));
    }
  }
  const int32_t wave = tid >> 6, lane = tid & 63;
  // Stage the box with direct global -> LDS loads (no registers, all in
  // flight at once): wave w copies rows w, w + 16, ..., each row in
  // segments of 64 lanes, lane l 16 bytes of a segment when the copy below
  // allows it, else one dword (coalesced along x, wrapped modulo the image
  // size per element or per 16-byte quad).
  {
    constexpr int32_t kDw = static_cast<int32_t>(sizeof(T) / 4);  // dwords per element
    constexpr int32_t kSegElems = 64 / kDw;                        // elements per segment
    const int32_t nseg = (bw + kSegElems - 1) / kSegElems;
    // one add or subtract wraps a box coordinate when the whole box lies
    // within one image size of the image (rows [-ny, 2 ny), columns alike);
    // otherwise (images smaller than the box) the general modulo
    const bool once_y = i_box0 >= -g.ny && int64_t{i_box0} + bh <= 2 * int64_t{g.ny};
    const bool once_x = j_box0 >= -g.nx && int64_t{j_box0} + bw <= 2 * int64_t{g.nx};
    // 16-byte copies (lane l fills bytes 16 l .. 16 l + 15 after the row's
    // LDS base): each lane moves kQuad consecutive elements.  With x 16-byte
    // aligned and nx, halo multiples of kQuad, every box column quad starts
    // at a source column c == 0 (mod kQuad), c + kQuad <= nx (the wrap points
    // 0 and nx fall between quads), and bw is a multiple of kQuad.
    constexpr int32_t kQuad = static_cast<int32_t>(16 / sizeof(T));
    const bool quads = (reinterpret_cast<uintptr_t>(x) & 15u) == 0 && g.nx % kQuad == 0 &&
                       halo % kQuad == 0;
    if (quads) {
      const int32_t nq = bw / kQuad;  // quads per box row
      for (int32_t a = __builtin_amdgcn_readfirstlane(wave); a < bh; a += kLdsThreads / 64) {
        int32_t r = i_box0 + a;
        r = once_y ? (r < 0 ? r + g.ny : (r >= g.ny ? r - g.ny : r)) : wrap_index(r, g.ny);
        const T* src_row = x + static_cast<int64_t>(r) * g.nx;
        for (int32_t q0 = 0; q0 < nq; q0 += 64) {
          if (q0 + lane < nq) {
            int32_t c = j_box0 + (q0 + lane) * kQuad;
            c = once_x ? (c < 0 ? c + g.nx : (c >= g.nx ? c - g.nx : c)) : wrap_index(c, g.nx);
            __builtin_amdgcn_global_load_lds(
                (const unsigned __attribute__((address_space(1)))*)(src_row + c),
                (unsigned __attribute__((address_space(3)))*)(box + (a * stride + q0 * kQuad) *
                                                                         static_cast<int32_t>(sizeof(T))),
                16, 0, 0);
          }
        }
      }
    } else
    for (int32_t a = __builtin_amdgcn_readfirstlane(wave); a < bh; a += kLdsThreads / 64) {
      int32_t r = i_box0 + a;
      r = once_y ? (r < 0 ? r + g.ny : (r >= g.ny ? r - g.ny : r)) : wrap_index(r, g.ny);
      const T* src_row = x + static_cast<int64_t>(r) * g.nx;
      for (int32_t sg = 0; sg < nseg; ++sg) {
        const int32_t b = sg * kSegElems + lane / kDw;  // element of the box row
        if (b < bw) {
          int32_t c = j_box0 + b;
          c = once_x ? (c < 0 ? c + g.nx : (c >= g.nx ? c - g.nx : c)) : wrap_index(c, g.nx);
          const unsigned char* src =
              reinterpret_cast<const unsigned char*>(src_row + c) + 4 * (lane % kDw);
          __builtin_amdgcn_global_load_lds(
              (const unsigned __attribute__((address_space(1)))*)src,
              (unsigned __attribute__((address_space(3)))*)(box + (a * stride + sg * kSegElems) *
                                                                       static_cast<int32_t>(sizeof(T))),
              4, 0, 0);
        }
      }
    }
    __builtin_amdgcn_s_waitcnt(0);  // this wave's copies have landed
  }
  __syncthreads();
  // wave w covers the 8 x 8 square at (8 * (w / 4), 8 * (w % 4)) of each
  // 32 x 32 quarter e of the tile
  const int32_t py0 = ((wave >> 2) << 3) | (lane >> 3);
  const int32_t px0 = ((wave & 3) << 3) | (lane & 7);
  const int32_t n_fwd = klen - 1 - kmid, n_bwd = kmid;
  const int64_t np = static_cast<int64_t>(g.ny) * g.nx;
  T acc[E];
  lds_byte* at0[E];
  bool inside[E];
  int64_t p[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int32_t py = py0 + 32 * (e / (TW / 32)), px = px0 + 32 * (e % (TW / 32));
    const int32_t i0 = ty * TH + py, j0 = tx * TW + px;
    inside[e] = i0 < g.ny && j0 < g.nx;
    p[e] = inside[e] ? static_cast<int64_t>(i0) * g.nx + j0 : 0;
    at0[e] = (lds_byte*)box + ((halo + py) * stride + (halo + px)) * elem;
    acc[e] = macc<FMA>(kernel[kmid], lds_value<T>(at0[e]), T(0));
  }
  // steps inside the box, then (HYBRID) the rest from the image: the box
  // position becomes an image position, wrapped (a periodic edge is just
  // crossed in the box); closed edges never leave the image (their clamps
  // were recorded as "stay")
  const int32_t ch_box = HYBRID ? halo / kChunkSteps : INT32_MAX;
  // one direction (fwd: std::true_type for forward) of every line of the lane
  auto direction = [&](auto fwd, int32_t nsteps, int32_t c0) {
    constexpr bool kForward = decltype(fwd)::value;
    lds_byte* at[E];
    bool alive[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      at[e] = at0[e];
      alive[e] = inside[e];
    }
    replay_lines<T, FMA, kForward, E>(LdsWalk<T>{}, acc, at, alive, kernel, kmid, nsteps,
                                      pair_delta, codes, np, p, c0, 0, ch_box);
    if constexpr (HYBRID) {
      if (nsteps > halo) {  // uniform
        // built here, next to its use, so that its descriptor stays in
        // scalar registers
        const ImageWalk<T> image(x, np);
        uint32_t pos[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int32_t cell = static_cast<int32_t>(at[e] - (lds_byte*)box) / elem;
          const int32_t r = cell / stride, c = cell - r * stride;
          const int32_t i = wrap_index(i_box0 + r, g.ny), j = wrap_index(j_box0 + c, g.nx);
          pos[e] = static_cast<uint32_t>(static_cast<int64_t>(i) * g.nx + j) * elem;
        }
        // float: every line of the lane at once (110 VGPRs, within the 128
        // of 4 waves per SIMD; 1-4% faster than two rounds of two);
        // double: two lines per lane at a time, half the registers of E
        constexpr int E2 = sizeof(T) == 4 ? E : (E >= 2 ? 2 : 1);
#pragma unroll
        for (int e0 = 0; e0 < E; e0 += E2) {
          T acc2[E2];
          uint32_t pos2[E2];
          bool alive2[E2];
          int64_t p2[E2];
#pragma unroll
          for (int k = 0; k < E2; ++k) {
            acc2[k] = acc[e0 + k];
            pos2[k] = pos[e0 + k];
            alive2[k] = alive[e0 + k];
            p2[k] = p[e0 + k];
          }
          replay_lines<T, FMA, kForward, E2>(image, acc2, pos2, alive2, kernel, kmid, nsteps,
                                             pair_image, codes, np, p2, c0, ch_box, INT32_MAX);
#pragma unroll
          for (int k = 0; k < E2; ++k) acc[e0 + k] = acc2[k];
        }
      }
    }
  };
  direction(std::true_type{}, n_fwd, 0);
  direction(std::false_type{}, n_bwd, chunks_of(n_fwd));
#pragma unroll
  for (int e = 0; e < E; ++e)
    if (inside[e]) out[p[e]] = acc[e];
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

TileMap make_tile_map(int32_t ny, int32_t nx, int32_t band, int32_t wave_w_log2) {
  TileMap m;
  m.wave_w_log2 = wave_w_log2 < 3 ? 3 : (wave_w_log2 > 5 ? 5 : wave_w_log2);
  const int32_t tw = tile_w(m), th = tile_h(m);
  m.tiles_x = (nx + tw - 1) / tw;
  m.tiles_y = (ny + th - 1) / th;
  m.n_tiles = m.tiles_x * m.tiles_y;
  m.per_xcd = m.n_tiles / kXcds;
  m.rem_xcd = m.n_tiles % kXcds;
  m.band = band < 1 ? 1 : band;
  return m;
}

int32_t code_chunks(int32_t klen) {
  const int32_t kmid = klen / 2;
  const int32_t n_fwd = klen - 1 - kmid, n_bwd = kmid;
  return chunks_of(n_fwd) + chunks_of(n_bwd);
}

template <typename T>
hipError_t launch_pack(const T* tex, int64_t tsy, int64_t tsx, const T* u, int64_t usy,
                       int64_t usx, const T* v, int64_t vsy, int64_t vsx, const FieldGeom& g,
                       Cell<T>* cells, hipStream_t stream, const T* kernel, int32_t klen,
                       T* kpad) {
  dim3 grid((g.nx + 2 + 63) / 64, (g.ny + 2 + 3) / 4);
  if (kpad) {  // enough threads for the weights of a long kernel on a small image
    const int64_t need = (int64_t{klen} + 2 * kKernelPad + 255) / 256;
    const int64_t have = static_cast<int64_t>(grid.x) * grid.y;
    if (need > have) grid.x = static_cast<unsigned>(std::min<int64_t>(
                         (need + grid.y - 1) / grid.y, std::max<int64_t>(grid.x, 1024)));
  }
  hipLaunchKernelGGL(lic_pack<T>, grid, dim3(256), 0, stream, tex, tsy, tsx, u, usy, usx, v, vsy,
                     vsx, g, cells, kernel, klen, kpad);
  return hipGetLastError();
}

int32_t kernel_pad() { return kKernelPad; }


template <typename T>
hipError_t launch_repack_texture(const T* x, const FieldGeom& g, Cell<T>* cells,
                                 hipStream_t stream) {
  if (static_cast<int64_t>(g.ny) * g.nx == 0) return hipSuccess;
  dim3 grid((g.nx + 2 + 63) / 64, (g.ny + 2 + 3) / 4);
  hipLaunchKernelGGL(lic_repack_texture<T>, grid, dim3(256), 0, stream, x, g, cells);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_copy2d(const T* src, int64_t sy, int64_t sx, int32_t ny, int32_t nx, T* dst,
                         hipStream_t stream) {
  dim3 grid((nx + 63) / 64, (ny + 3) / 4);
  hipLaunchKernelGGL(lic_copy2d<T>, grid, dim3(256), 0, stream, src, sy, sx, ny, nx, dst);
  return hipGetLastError();
}

// Buffer addressing covers arrays of fewer than 2^32 bytes.
inline bool fits_buffer(int64_t bytes) { return bytes > 0 && bytes <= 0xFFFFFFFFll; }


template <typename T, bool FMA, bool BL, bool POL, bool REC>
static hipError_t trace_one(const Cell<T>* cells, const T* kernel, int32_t klen,
                            const FieldGeom& g, const TileMap& m, T* out, uint4* codes,
                            hipStream_t stream) {
  if (fits_buffer(guarded_cells(g.ny, g.nx) * static_cast<int64_t>(sizeof(Cell<T>)))) {
    hipLaunchKernelGGL((lic_trace<T, FMA, BL, POL, REC, true>), dim3(m.n_tiles), dim3(kBlock),
                       0, stream, cells, kernel, klen, g, m, out, codes);
  } else {
    hipLaunchKernelGGL((lic_trace<T, FMA, BL, POL, REC, false>), dim3(m.n_tiles), dim3(kBlock),
                       0, stream, cells, kernel, klen, g, m, out, codes);
  }
  return hipGetLastError();
}

template <typename T, bool FMA, bool BL>
static hipError_t trace_mode(const Cell<T>* cells, const T* kernel, int32_t klen,
                             const FieldGeom& g, const TileMap& m, int uv_mode, T* out,
                             uint4* codes, hipStream_t stream) {
  const bool rec = codes != nullptr;
  if (uv_mode == 1) {
    return rec ? trace_one<T, FMA, BL, true, true>(cells, kernel, klen, g, m, out, codes, stream)
               : trace_one<T, FMA, BL, true, false>(cells, kernel, klen, g, m, out, codes,
                                                    stream);
  }
  return rec ? trace_one<T, FMA, BL, false, true>(cells, kernel, klen, g, m, out, codes, stream)
             : trace_one<T, FMA, BL, false, false>(cells, kernel, klen, g, m, out, codes, stream);
}

template <typename T>
hipError_t launch_trace(const Cell<T>* cells, const T* kernel, int32_t klen, const FieldGeom& g,
                        const TileMap& m, int uv_mode, int profile, T* out, uint4* codes,
                        hipStream_t stream) {
  switch (profile & 3) {
    case 0:
      return trace_mode<T, false, false>(cells, kernel, klen, g, m, uv_mode, out, codes, stream);
    case 1:
      return trace_mode<T, true, false>(cells, kernel, klen, g, m, uv_mode, out, codes, stream);
    case 2:
      return trace_mode<T, false, true>(cells, kernel, klen, g, m, uv_mode, out, codes, stream);
    default:
      return trace_mode<T, true, true>(cells, kernel, klen, g, m, uv_mode, out, codes, stream);
  }
}

size_t trace_pair_lds_bytes(int32_t klen, size_t elem) {
  return size_t{32} * static_cast<size_t>(klen / 2 + 1) * elem;
}

template <typename T, bool FMA, bool BL, bool POL, bool REC>
static hipError_t trace_pair_one(const Cell<T>* cells, const T* kernel, int32_t klen,
                                 const FieldGeom& g, const TileMap& m, T* out, uint4* codes,
                                 hipStream_t stream) {
  const size_t lds = trace_pair_lds_bytes(klen, sizeof(T));
  auto* fn = &lic_trace_pair<T, FMA, BL, POL, REC>;
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                            hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
  hipLaunchKernelGGL(fn, dim3(m.n_tiles * 8), dim3(64), lds, stream, cells, kernel, klen, g, m,
                     out, codes);
  return hipGetLastError();
}

template <typename T, bool FMA, bool BL>
static hipError_t trace_pair_mode(const Cell<T>* cells, const T* kernel, int32_t klen,
                                  const FieldGeom& g, const TileMap& m, int uv_mode, T* out,
                                  uint4* codes, hipStream_t stream) {
  const bool rec = codes != nullptr;
  if (uv_mode == 1) {
    return rec ? trace_pair_one<T, FMA, BL, true, true>(cells, kernel, klen, g, m, out, codes,
                                                         stream)
               : trace_pair_one<T, FMA, BL, true, false>(cells, kernel, klen, g, m, out, codes,
                                                          stream);
  }
  return rec ? trace_pair_one<T, FMA, BL, false, true>(cells, kernel, klen, g, m, out, codes,
                                                        stream)
             : trace_pair_one<T, FMA, BL, false, false>(cells, kernel, klen, g, m, out, codes,
                                                         stream);
}

template <typename T>
hipError_t launch_replay(const T* x, const T* kernel, int32_t klen, const FieldGeom& g,
                         const TileMap& m, int profile, T* out, const uint4* codes,
                         hipStream_t stream) {
  const bool buf = fits_buffer(static_cast<int64_t>(g.ny) * g.nx * sizeof(T));
  if (profile & 1) {
    if (buf) {
      hipLaunchKernelGGL((lic_replay<T, true, true>), dim3(m.n_tiles), dim3(kBlock), 0, stream,
                         x, kernel, klen, g, m, out, codes);
    } else {
      hipLaunchKernelGGL((lic_replay<T, true, false>), dim3(m.n_tiles), dim3(kBlock), 0, stream,
                         x, kernel, klen, g, m, out, codes);
    }
  } else {
    if (buf) {
      hipLaunchKernelGGL((lic_replay<T, false, true>), dim3(m.n_tiles), dim3(kBlock), 0, stream,
                         x, kernel, klen, g, m, out, codes);
    } else {
      hipLaunchKernelGGL((lic_replay<T, false, false>), dim3(m.n_tiles), dim3(kBlock), 0,
                         stream, x, kernel, klen, g, m, out, codes);
    }
  }
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void lic_div_check(const float* __restrict__ a,
                                                     const float* __restrict__ b,
                                                     float* __restrict__ q, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) q[i] = div_rn(a[i], b[i]);
}

hipError_t launch_div_check(const float* a, const float* b, float* q, int64_t n,
                            hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(lic_div_check, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                     stream, a, b, q, n);
  return hipGetLastError();
}

hipError_t launch_record_stats(const uint4* codes, int32_t klen, const FieldGeom& g,
                               const TileMap& m, unsigned long long* stats, hipStream_t stream) {
  hipLaunchKernelGGL(lic_record_stats, dim3(m.n_tiles), dim3(kBlock), 0, stream, codes, klen, g,
                     m, stats);
  return hipGetLastError();
}

int32_t lds_replay_limit_bytes() {
  int dev = 0, limit = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return 0;
  return limit;
}

template <typename T>
bool trace_pair_fits(int32_t klen, int64_t ny, int64_t nx) {
  const size_t limit = static_cast<size_t>(lds_replay_limit_bytes());
  return klen % 2 == 1 && trace_pair_lds_bytes(klen, sizeof(T)) <= limit &&
         fits_buffer(guarded_cells(ny, nx) * static_cast<int64_t>(sizeof(Cell<T>)));
}

template <typename T>
hipError_t launch_trace_pair(const Cell<T>* cells, const T* kernel, int32_t klen,
                             const FieldGeom& g, const TileMap& m, int uv_mode, int profile,
                             T* out, uint4* codes, hipStream_t stream) {
  if (!trace_pair_fits<T>(klen, g.ny, g.nx)) return hipErrorInvalidValue;
  switch (profile & 3) {
    case 0:
      return trace_pair_mode<T, false, false>(cells, kernel, klen, g, m, uv_mode, out, codes,
                                              stream);
    case 1:
      return trace_pair_mode<T, true, false>(cells, kernel, klen, g, m, uv_mode, out, codes,
                                             stream);
    case 2:
      return trace_pair_mode<T, false, true>(cells, kernel, klen, g, m, uv_mode, out, codes,
                                             stream);
    default:
      return trace_pair_mode<T, true, true>(cells, kernel, klen, g, m, uv_mode, out, codes,
                                            stream);
  }
}

// LDS replay tile for this kernel length: the largest of 64x64, 32x64 and
// 32x32 whose staged box fits the per-block LDS; {0, 0} when none does.
template <typename T>
LdsTile lds_replay_tile(int32_t klen) {
  const size_t limit = static_cast<size_t>(lds_replay_limit_bytes());
  const int32_t shapes[3][2] = {{64, 64}, {32, 64}, {32, 32}};
  for (const auto& sh : shapes) {
    if (lds_replay_bytes(sh[0], sh[1], klen / 2, sizeof(T)) <= limit) return {sh[0], sh[1]};
  }
  return {0, 0};
}

template <typename T>
bool lds_replay_fits(int32_t klen) {
  return lds_replay_tile<T>(klen).h > 0;
}

template <typename T>
bool lds_hybrid_fits(int32_t klen, int64_t n, int32_t halo) {
  return halo > 0 && halo % kChunkSteps == 0 && halo <= hybrid_max_halo<T>() &&
         klen / 2 > halo && fits_buffer(n * static_cast<int64_t>(sizeof(T))) &&
         lds_replay_bytes(64, 64, halo, sizeof(T), 2) <=
             static_cast<size_t>(lds_replay_limit_bytes());
}

template <typename T, bool FMA, int TH, int TW, bool HYBRID = false>
static void replay_lds_one(const T* x, const T* kernel, int32_t klen, const FieldGeom& g, T* out,
                           const uint4* codes, hipStream_t stream, int32_t hybrid_halo = 0) {
  TileMap m;
  m.wave_w_log2 = 3;
  m.tiles_x = (g.nx + TW - 1) / TW;
  m.tiles_y = (g.ny + TH - 1) / TH;
  m.n_tiles = m.tiles_x * m.tiles_y;
  m.per_xcd = m.n_tiles / kXcds;
  m.rem_xcd = m.n_tiles % kXcds;
  m.band = 4;
  const int32_t halo = HYBRID ? hybrid_halo : klen / 2;
  const size_t bytes = lds_replay_dynamic_bytes(TH, TW, halo, sizeof(T));
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&lic_replay_lds<T, FMA, TH, TW, HYBRID>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  (void)hipGetLastError();  // the attribute is advisory on this target
  hipLaunchKernelGGL((lic_replay_lds<T, FMA, TH, TW, HYBRID>), dim3(m.n_tiles), dim3(kLdsThreads),
                     bytes, stream, x, kernel, klen, g, m, out, codes, halo);
}

template <typename T, bool FMA>
static void replay_lds_shape(const T* x, const T* kernel, int32_t klen, const FieldGeom& g,
                             T* out, const uint4* codes, hipStream_t stream) {
  const LdsTile t = lds_replay_tile<T>(klen);
  if (t.h == 64) {
    replay_lds_one<T, FMA, 64, 64>(x, kernel, klen, g, out, codes, stream);
  } else if (t.w == 64) {
    replay_lds_one<T, FMA, 32, 64>(x, kernel, klen, g, out, codes, stream);
  } else {
    replay_lds_one<T, FMA, 32, 32>(x, kernel, klen, g, out, codes, stream);
  }
}

template <typename T>
hipError_t launch_replay_lds(const T* x, const T* kernel, int32_t klen, const FieldGeom& g,
                             int profile, T* out, const uint4* codes, hipStream_t stream) {
  if (profile & 1) {
    replay_lds_shape<T, true>(x, kernel, klen, g, out, codes, stream);
  } else {
    replay_lds_shape<T, false>(x, kernel, klen, g, out, codes, stream);
  }
  return hipGetLastError();
}

template <typename T, bool FMA>
static bool replay_lds_tile_fma(const T* x, const T* kernel, int32_t klen, const FieldGeom& g,
                                LdsTile tile, T* out, const uint4* codes, hipStream_t stream) {
  if (tile.h == 64 && tile.w == 64) {
    replay_lds_one<T, FMA, 64, 64>(x, kernel, klen, g, out, codes, stream);
  } else if (tile.h == 32 && tile.w == 64) {
    replay_lds_one<T, FMA, 32, 64>(x, kernel, klen, g, out, codes, stream);
  } else if (tile.h == 32 && tile.w == 32) {
    replay_lds_one<T, FMA, 32, 32>(x, kernel, klen, g, out, codes, stream);
  } else {
    return false;
  }
  return true;
}

template <typename T>
hipError_t launch_replay_lds_tile(const T* x, const T* kernel, int32_t klen, const FieldGeom& g,
                                  LdsTile tile, int profile, T* out, const uint4* codes,
                                  hipStream_t stream) {
  if (lds_replay_bytes(tile.h, tile.w, klen / 2, sizeof(T)) >
      static_cast<size_t>(lds_replay_limit_bytes()))
    return hipErrorInvalidValue;
  const bool ok = (profile & 1)
                      ? replay_lds_tile_fma<T, true>(x, kernel, klen, g, tile, out, codes, stream)
                      : replay_lds_tile_fma<T, false>(x, kernel, klen, g, tile, out, codes, stream);
  return ok ? hipGetLastError() : hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_replay_lds_hybrid(const T* x, const T* kernel, int32_t klen, const FieldGeom& g,
                                    int32_t halo, int profile, T* out, const uint4* codes,
                                    hipStream_t stream) {
  if (!lds_hybrid_fits<T>(klen, static_cast<int64_t>(g.ny) * g.nx, halo))
    return hipErrorInvalidValue;
  if (profile & 1) {
    replay_lds_one<T, true, 64, 64, true>(x, kernel, klen, g, out, codes, stream, halo);
  } else {
    replay_lds_one<T, false, 64, 64, true>(x, kernel, klen, g, out, codes, stream, halo);
  }
  return hipGetLastError();
}

#define RLIC_INSTANTIATE(T)                                                                     \
  template hipError_t launch_pack<T>(const T*, int64_t, int64_t, const T*, int64_t, int64_t,   \
                                     const T*, int64_t, int64_t, const FieldGeom&, Cell<T>*,   \
                                     hipStream_t, const T*, int32_t, T*);                      \
  template hipError_t launch_repack_texture<T>(const T*, const FieldGeom&, Cell<T>*,           \
                                               hipStream_t);                                    \
  template hipError_t launch_copy2d<T>(const T*, int64_t, int64_t, int32_t, int32_t, T*,       \
                                       hipStream_t);                                           \
  template hipError_t launch_trace<T>(const Cell<T>*, const T*, int32_t, const FieldGeom&,     \
                                      const TileMap&, int, int, T*, uint4*, hipStream_t);   \
  template hipError_t launch_replay<T>(const T*, const T*, int32_t, const FieldGeom&,          \
                                       const TileMap&, int, T*, const uint4*, hipStream_t);    \
  template bool lds_replay_fits<T>(int32_t);                                                   \
  template bool trace_pair_fits<T>(int32_t, int64_t, int64_t);                               \
  template hipError_t launch_trace_pair<T>(const Cell<T>*, const T*, int32_t, const FieldGeom&, \
                                           const TileMap&, int, int, T*, uint4*, hipStream_t);  \
  template LdsTile lds_replay_tile<T>(int32_t);                                                \
  template hipError_t launch_replay_lds<T>(const T*, const T*, int32_t, const FieldGeom&, int, \
                                           T*, const uint4*, hipStream_t);                      \
  template hipError_t launch_replay_lds_tile<T>(const T*, const T*, int32_t, const FieldGeom&,  \
                                                LdsTile, int, T*, const uint4*, hipStream_t);  \
  template bool lds_hybrid_fits<T>(int32_t, int64_t, int32_t);                                 \
  template hipError_t launch_replay_lds_hybrid<T>(const T*, const T*, int32_t,                 \
                                                  const FieldGeom&, int32_t, int, T*,          \
                                                  const uint4*, hipStream_t);

RLIC_INSTANTIATE(float)
RLIC_INSTANTIATE(double)

}  // namespace rlic
