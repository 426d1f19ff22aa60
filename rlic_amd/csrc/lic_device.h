// Device-side building blocks of the streamline kernels (gfx950 / CDNA4).
//
// Semantics follow the reference's per-step rule (src/lib.rs:157-273,
// 304-362) exactly; the arithmetic is literal so that results are
// bit-identical to the reference's CPU arithmetic:
//   * fused multiply-add only where the reference calls `mul_add`
//     (src/lib.rs:177, 231, 359, 382), selected by the FMA template flag;
//   * every other product is a separate rounded multiply (the library is
//     compiled with -ffp-contract=off);
//   * divisions are IEEE correctly rounded (hipcc's default fp32 division
//     and the fp64 div_scale/div_fmas/div_fixup sequence), denormals kept.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rlic {

// One packed field cell: the vector (u, v) at the cell and the texture
// value x.  A streamline step reads the texture at the cell it enters and,
// on the next step, the vector at that same cell, so packing the three
// values turns three gathers per step into one 16-byte (fp32) or two
// adjacent 16-byte (fp64) loads from a single cache line.
template <typename T>
struct Cell;

template <>
struct __align__(16) Cell<float> {
  float u, v, x, pad;
};

template <>
struct __align__(32) Cell<double> {
  double u, v, x, pad;
};

template <typename T>
struct CellVal {
  T u, v, x;
  T pad;  // a guarded array's jump (see pad_bits)
};

// The jump in cells that a guarded cell array keeps in a cell's pad slot
// (low 32 bits of its bit pattern; 0 in image cells).
__device__ __forceinline__ uint32_t pad_bits(float pad) { return __builtin_bit_cast(uint32_t, pad); }
__device__ __forceinline__ uint32_t pad_bits(double pad) {
  return static_cast<uint32_t>(__builtin_bit_cast(uint64_t, pad));
}
__host__ __device__ inline float jump_pad(int32_t jump, float*) {
  return __builtin_bit_cast(float, jump);
}
__host__ __device__ inline double jump_pad(int32_t jump, double*) {
  return __builtin_bit_cast(double, static_cast<uint64_t>(static_cast<uint32_t>(jump)));
}

__device__ __forceinline__ CellVal<float> load_cell(const Cell<float>* __restrict__ p,
                                                    uint32_t idx) {
  const float4 c = *reinterpret_cast<const float4*>(p + idx);
  return {c.x, c.y, c.z, c.w};
}

__device__ __forceinline__ CellVal<double> load_cell(const Cell<double>* __restrict__ p,
                                                     uint32_t idx) {
  const double2* q = reinterpret_cast<const double2*>(p + idx);
  const double2 a = q[0];
  const double2 b = q[1];
  return {a.x, a.y, b.x, b.y};
}

// ---------------------------------------------------------------------------
// Gathers.  With BUF, arrays are addressed through a raw buffer resource and
// a 32-bit byte offset (one VALU add per address, hardware range check: an
// out-of-range read returns 0 instead of faulting); usable when the array
// spans fewer than 2^32 bytes.  Otherwise 64-bit global addresses are used.
// ---------------------------------------------------------------------------
// gfx9 raw-buffer descriptor word 3 (no swizzle, dword data format).
constexpr int32_t kBufferWord3 = 0x00020000;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                           kBufferWord3);
}

__device__ __forceinline__ float buf_load(const __amdgpu_buffer_rsrc_t r, uint32_t off, float*) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ double buf_load(const __amdgpu_buffer_rsrc_t r, uint32_t off,
                                           double*) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

// The b64/b128 builtins return GCC vector_size vectors of unsigned int;
// they are bit-cast whole (bit-casting a subscripted lane miscompiles).
__device__ __forceinline__ CellVal<float> buf_load_cell(const __amdgpu_buffer_rsrc_t r,
                                                        uint32_t off, float*) {
  const float4 c = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
  return {c.x, c.y, c.z, c.w};
}
__device__ __forceinline__ CellVal<double> buf_load_cell(const __amdgpu_buffer_rsrc_t r,
                                                         uint32_t off, double*) {
  const double2 uv =
      __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
  const double2 xp =
      __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off + 16u, 0, 0));
  return {uv.x, uv.y, xp.x, xp.y};
}

// Cell gather at `pos`: a byte offset (BUF) or an element index (global).
template <typename T, bool BUF>
struct CellGather {
  const Cell<T>* base;
  __amdgpu_buffer_rsrc_t rsrc;
  __device__ __forceinline__ CellGather(const Cell<T>* p, int64_t n) : base(p) {
    if constexpr (BUF) rsrc = make_rsrc(p, static_cast<uint32_t>(n * sizeof(Cell<T>)));
  }
  static constexpr uint32_t unit = BUF ? static_cast<uint32_t>(sizeof(Cell<T>)) : 1u;
  __device__ __forceinline__ CellVal<T> operator()(uint32_t pos) const {
    if constexpr (BUF) {
      return buf_load_cell(rsrc, pos, static_cast<T*>(nullptr));
    } else {
      return load_cell(base, pos);
    }
  }
};

// Texture gather at `pos`: a byte offset (BUF) or an element index (global).
template <typename T, bool BUF>
struct ValueGather {
  const T* base;
  __amdgpu_buffer_rsrc_t rsrc;
  __device__ __forceinline__ ValueGather(const T* p, int64_t n) : base(p) {
    if constexpr (BUF) rsrc = make_rsrc(p, static_cast<uint32_t>(n * sizeof(T)));
  }
  static constexpr uint32_t unit = BUF ? static_cast<uint32_t>(sizeof(T)) : 1u;
  __device__ __forceinline__ T operator()(uint32_t pos) const {
    if constexpr (BUF) {
      return buf_load(rsrc, pos, static_cast<T*>(nullptr));
    } else {
      return base[pos];
    }
  }
};

// A copy of a (wave-uniform) value that the compiler must treat as
// lane-varying: loop-invariant select operands kept this way live in VGPRs,
// so each select is one v_cndmask instead of a scalar-to-vector move (a
// VALU instruction reads at most one scalar register) or a branch.
__device__ __forceinline__ uint32_t lane_copy(uint32_t x) {
  uint32_t r;
  asm("v_mov_b32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// x with its sign bit xor-ed with `sign` (0 or 0x80000000; for double it
// applies to the top word): -x when the bit is set, bit for bit the
// negation, as one integer instruction.
__device__ __forceinline__ float xor_sign(float x, uint32_t sign) {
  return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, x) ^ sign);
}
__device__ __forceinline__ double xor_sign(double x, uint32_t sign) {
  return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, x) ^ (uint64_t{sign} << 32));
}

// 2 * x + (this lane's bit of the wave64 lane mask `mask`) in one VALU
// instruction: the mask is the add's carry-in.
__device__ __forceinline__ uint32_t carry_in_bit(uint32_t x, uint64_t mask) {
  uint32_t r;
  uint64_t carry_out;
  asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(r), "=s"(carry_out) : "v"(x), "s"(mask));
  return r;
}

// This lane's bit of the wave64 lane mask `mask`, as 0 or 1.
__device__ __forceinline__ uint32_t lane_bit(uint64_t mask) {
  uint32_t r;
  uint64_t carry_out;
  asm("v_addc_co_u32_e64 %0, %1, 0, 0, %2" : "=v"(r), "=s"(carry_out) : "s"(mask));
  return r;
}

__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float abs_t(float a) { return __builtin_fabsf(a); }
__device__ __forceinline__ double abs_t(double a) { return __builtin_fabs(a); }
__device__ __forceinline__ bool isnan_t(float a) { return __builtin_isnan(a); }
__device__ __forceinline__ bool isnan_t(double a) { return __builtin_isnan(a); }

template <typename T>
__device__ __forceinline__ T inf_t() {
  return static_cast<T>(__builtin_huge_valf());
}

// `a.mul_add(b, c)` with the fma feature, else `c += a * b` (two roundings).
template <bool FMA, typename T>
__device__ __forceinline__ T macc(T a, T b, T c) {
  if constexpr (FMA) {
    return fma_t(a, b, c);
  } else {
    const T prod = a * b;
    return c + prod;
  }
}

// Time to reach the next cell edge along one axis (src/lib.rs:157-180).
// Correctly rounded division (IEEE round-to-nearest-even, denormals kept):
// hipcc's float and double division (v_div_scale / v_rcp / v_div_fmas /
// v_div_fixup).  A float quotient formed in double from a Newton-refined
// reciprocal and rounded once is also exact (8 instructions instead of 11)
// but measured slower on gfx950 (trace 3.54 ms vs 3.10 ms at 4096^2, L=129):
// the double-precision operations cost more issue cycles.
__device__ __forceinline__ float div_rn(float a, float b) { return a / b; }
__device__ __forceinline__ double div_rn(double a, double b) { return a / b; }

// `velocity` is never NaN here: NaN vectors end the streamline first
// (src/lib.rs:336-338).
// `negative` is the sign bit of `velocity` (the branchless rule's
// 1 + signum(velocity)), given by the caller, which tests it once for
// both uses.
template <typename T, bool FMA, bool BRANCHLESS>
__device__ __forceinline__ T time_to_next_pixel(T velocity, T frac, bool negative) {
  if constexpr (!BRANCHLESS) {
    if (velocity > T(0)) return div_rn(T(1) - frac, velocity);
    if (velocity < T(0)) return -div_rn(frac, velocity);
    return inf_t<T>();
  } else {
    // 1 + signum(velocity): 0 when the sign bit is set (also -0.0), else 2.
    const T factor = negative ? T(0) : T(2);
    const T remaining = macc<FMA>(factor, T(0.5) - frac, frac);
    return abs_t(div_rn(remaining, velocity));
  }
}

template <typename T, bool FMA, bool BRANCHLESS>
__device__ __forceinline__ T time_to_next_pixel(T velocity, T frac) {
  return time_to_next_pixel<T, FMA, BRANCHLESS>(velocity, frac, __builtin_signbit(velocity));
}

struct FieldGeom {
  int32_t ny, nx;
  int32_t x_left, x_right, y_left, y_right;  // 1 = periodic, 0 = closed
};

// Per-field constant of the incremental step rule on a guarded cell array
// (positions in `unit`s): the row step.
struct StepGeom {
  uint32_t grow;
};

// Cells of the guarded array of an ny x nx field: (ny + 2) x (nx + 2), the
// image's cell (i, j) at (i + 1, j + 1), a ring of guard cells around it.
__host__ __device__ inline int64_t guarded_cells(int64_t ny, int64_t nx) {
  return (ny + 2) * (nx + 2);
}

// Image cell (ti, tj) whose values guard cell (gi, gj) of the guarded array
// holds: itself for an image cell, else the cell the boundary rule sends a
// step onto the guard cell to (src/lib.rs:83-95: the edge cell on a closed
// side, the opposite edge cell on a periodic one).  False for the four
// corners, which no step reaches (a step changes one coordinate).
__host__ __device__ inline bool guard_source(int32_t gi, int32_t gj, int32_t ny, int32_t nx,
                                             int32_t x_left, int32_t x_right, int32_t y_left,
                                             int32_t y_right, int32_t& ti, int32_t& tj) {
  ti = gi - 1;
  tj = gj - 1;
  const bool out_y = ti < 0 || ti >= ny, out_x = tj < 0 || tj >= nx;
  if (out_y && out_x) return false;
  if (tj < 0) tj = x_left ? nx - 1 : 0;
  if (tj >= nx) tj = x_right ? 0 : nx - 1;
  if (ti < 0) ti = y_left ? ny - 1 : 0;
  if (ti >= ny) ti = y_right ? 0 : ny - 1;
  return true;
}

// Cell (gi, gj) of a field's guarded array from its strided texture, u, v.
template <typename T>
__device__ __forceinline__ Cell<T> guarded_cell(const T* __restrict__ tex, int64_t tsy,
                                                int64_t tsx, const T* __restrict__ u,
                                                int64_t usy, int64_t usx,
                                                const T* __restrict__ v, int64_t vsy,
                                                int64_t vsx, const FieldGeom& g, int32_t gi,
                                                int32_t gj) {
  Cell<T> c;
  int32_t i, j;
  if (guard_source(gi, gj, g.ny, g.nx, g.x_left, g.x_right, g.y_left, g.y_right, i, j)) {
    c.u = u[i * usy + j * usx];
    c.v = v[i * vsy + j * vsx];
    c.x = tex[i * tsy + j * tsx];
    c.pad = jump_pad((i + 1 - gi) * (g.nx + 2) + (j + 1 - gj), static_cast<T*>(nullptr));
  } else {
    c.u = T(0);
    c.v = T(0);
    c.x = T(0);
    c.pad = T(0);
  }
  return c;
}

// Move codes recorded by the tracing pass (4 bits per step, 8 per word), a
// bit field: kMoveMoved for a step to a neighbouring cell, plus kMoveAxisY
// for a step along y (rows), kMoveNeg toward lower indices and kMoveEdge
// when the step left the image (the boundary rule of that side then
// applies); kMoveStay for no move and kMoveStop where the streamline ended
// on a NaN vector (src/lib.rs:336-338).  Codes 1-6 do not occur.
enum : uint32_t {
  kMoveStay = 0,
  kMoveNeg = 1,
  kMoveAxisY = 2,
  kMoveEdge = 4,
  kMoveStop = 7,
  kMoveMoved = 8,
  kMoveXPlus = kMoveMoved,
  kMoveXMinus = kMoveMoved | kMoveNeg,
  kMoveYPlus = kMoveMoved | kMoveAxisY,
  kMoveYMinus = kMoveMoved | kMoveAxisY | kMoveNeg
};

// A word holds the codes of 8 consecutive steps, the first step in the top
// nibble: step q in bits 28 - 4q .. 31 - 4q.  (The tracing pass shifts
// each code bit in at the bottom of the word with one add-with-carry, so
// the first step's bits end on top.)
__host__ __device__ inline uint32_t move_code(uint32_t word, int q) {
  return (word >> (28 - 4 * q)) & 15u;
}

// Bit 4k set where nibble k of `word` is the stop code.
__host__ __device__ inline uint32_t stop_nibbles(uint32_t word) {
  return ~(word >> 3) & (word >> 2) & (word >> 1) & word & 0x11111111u;
}

// Steps of a word before its first stop code (8 when it has none): with
// the top stop bit at 4k (step 7 - k), (stops << 3) | 1 has its top bit at
// 4k + 3 (bit 0 without a stop), so (leading zeros + 1) / 4 = 7 - k (8);
// branch-free, and the argument of clz is never 0.
__host__ __device__ inline int32_t steps_before_stop(uint32_t word) {
  const uint32_t stops = stop_nibbles(word);
  return (__builtin_clz((stops << 3) | 1u) + 1) >> 2;
}

// Whether a word of 8 recorded steps holds a stop code.  The tracing pass
// records the stop code again at every step after a line has ended (the
// line stays on its NaN cell), so a word whose 8 steps were all recorded
// has one exactly when its last code (the bottom nibble) is the stop code.
__host__ __device__ inline bool group_has_stop(uint32_t word) {
  return (word & 15u) == kMoveStop;
}

// Positions are in `unit`s: 1 for element indices, the cell size for byte
// offsets.
__host__ __device__ inline StepGeom make_step_geom(const FieldGeom& g, uint32_t unit) {
  StepGeom s;
  s.grow = static_cast<uint32_t>(g.nx + 2) * unit;
  return s;
}

// Position change of each move code in `unit`s, added modulo 2^32
// (positions are uint32_t).  `row` is the row length and `span_x`, `span_y`
// the distance between opposite edge cells, all in units; a periodic edge
// step jumps across the image, a closed one stays.
__host__ __device__ inline uint32_t move_delta_units(uint32_t code, const FieldGeom& g,
                                                     uint32_t unit, uint32_t row,
                                                     uint32_t span_x, uint32_t span_y) {
  switch (code) {
    case kMoveXPlus: return unit;
    case kMoveXMinus: return 0u - unit;
    case kMoveYPlus: return row;
    case kMoveYMinus: return 0u - row;
    case kMoveXPlus | kMoveEdge: return g.x_right ? 0u - span_x : 0u;
    case kMoveXMinus | kMoveEdge: return g.x_left ? span_x : 0u;
    case kMoveYPlus | kMoveEdge: return g.y_right ? 0u - span_y : 0u;
    case kMoveYMinus | kMoveEdge: return g.y_left ? span_y : 0u;
    default: return 0u;
  }
}

// Replay through the image itself: wraps jump across it.
__host__ __device__ inline uint32_t move_delta(uint32_t code, const FieldGeom& g, uint32_t unit) {
  const uint32_t row = static_cast<uint32_t>(g.nx) * unit;
  return move_delta_units(code, g, unit, row, static_cast<uint32_t>(g.nx - 1) * unit,
                          static_cast<uint32_t>(g.ny - 1) * row);
}

// Block -> pixel tile.  A block is 4 wave64s in a 2x2 arrangement; each wave
// covers a wave_w x (64 / wave_w) rectangle, so a block tile is
// (2 * wave_w) x (128 / wave_w).  Workgroups are dealt round-robin to the
// 8 XCDs; each XCD is given a contiguous run of logical tiles, and logical
// tiles run column-major inside horizontal strips `band` tiles tall, so the
// tiles resident on one XCD at a time form a compact patch whose streamline
// halo stays in that XCD's L2.
struct TileMap {
  int32_t tiles_x, tiles_y, n_tiles;
  int32_t per_xcd, rem_xcd;  // n_tiles = 8 * per_xcd + rem_xcd
  int32_t band;
  int32_t wave_w_log2;       // 3, 4 or 5 (wave tile 8x8, 16x4, 32x2)
};

constexpr int kBlock = 256;
constexpr int kXcds = 8;

__host__ __device__ inline int32_t tile_w(const TileMap& m) { return 2 << m.wave_w_log2; }
__host__ __device__ inline int32_t tile_h(const TileMap& m) { return 2 << (6 - m.wave_w_log2); }

__device__ __forceinline__ void tile_of_block(int32_t b, const TileMap& m, int32_t& ty,
                                              int32_t& tx) {
  const int32_t xcd = b & (kXcds - 1);
  const int32_t slot = b >> 3;
  const int32_t logical = xcd * m.per_xcd + min(xcd, m.rem_xcd) + slot;
  const int32_t strip_tiles = m.band * m.tiles_x;
  const int32_t strip = logical / strip_tiles;
  const int32_t within = logical - strip * strip_tiles;
  const int32_t h = min(m.band, m.tiles_y - strip * m.band);
  tx = within / h;
  ty = strip * m.band + (within - tx * h);
}

// Thread -> pixel (py, px) inside the block tile.
__device__ __forceinline__ void pixel_of_thread(int32_t tid, const TileMap& m, int32_t& py,
                                                int32_t& px) {
  const int32_t w = tid >> 6, l = tid & 63;
  const int32_t ww = m.wave_w_log2, wh = 6 - m.wave_w_log2;
  px = ((w & 1) << ww) | (l & ((1 << ww) - 1));
  py = ((w >> 1) << wh) | (l >> ww);
}

}  // namespace rlic
