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
};

__device__ __forceinline__ CellVal<float> load_cell(const Cell<float>* __restrict__ p,
                                                    int32_t idx) {
  const float4 c = *reinterpret_cast<const float4*>(p + idx);
  return {c.x, c.y, c.z};
}

__device__ __forceinline__ CellVal<double> load_cell(const Cell<double>* __restrict__ p,
                                                     int32_t idx) {
  const double2* q = reinterpret_cast<const double2*>(p + idx);
  const double2 a = q[0];
  const double2 b = q[1];
  return {a.x, a.y, b.x};
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
// `velocity` is never NaN here: NaN vectors end the streamline first
// (src/lib.rs:336-338).
template <typename T, bool FMA, bool BRANCHLESS>
__device__ __forceinline__ T time_to_next_pixel(T velocity, T frac) {
  if constexpr (!BRANCHLESS) {
    if (velocity > T(0)) return (T(1) - frac) / velocity;
    if (velocity < T(0)) return -(frac / velocity);
    return inf_t<T>();
  } else {
    // 1 + signum(velocity): 0 when the sign bit is set (also -0.0), else 2.
    const T factor = __builtin_signbit(velocity) ? T(0) : T(2);
    const T remaining = macc<FMA>(factor, T(0.5) - frac, frac);
    return abs_t(remaining / velocity);
  }
}

// Boundary rule applied to one coordinate after each step
// (src/lib.rs:83-95): -1 is the "stepped off the low edge" sentinel.
__device__ __forceinline__ int32_t apply_bound(int32_t c, int32_t size, int32_t low_periodic,
                                               int32_t high_periodic) {
  if (c == -1) return low_periodic ? size - 1 : 0;
  if (c == size) return high_periodic ? 0 : size - 1;
  return c;
}

struct FieldGeom {
  int32_t ny, nx;
  int32_t x_left, x_right, y_left, y_right;  // 1 = periodic, 0 = closed
};

// Block -> 16x16 pixel tile.  Workgroups are dealt round-robin to the 8
// XCDs; each XCD is given a contiguous run of logical tiles, and logical
// tiles run column-major inside horizontal strips `band` tiles tall, so the
// tiles resident on one XCD at a time form a compact square whose
// streamline halo stays in that XCD's L2.
struct TileMap {
  int32_t tiles_x, tiles_y, n_tiles;
  int32_t per_xcd, rem_xcd;  // n_tiles = 8 * per_xcd + rem_xcd
  int32_t band;
};

constexpr int kTile = 16;
constexpr int kBlock = kTile * kTile;
constexpr int kXcds = 8;

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

// Thread -> pixel inside the tile: each wave64 covers an 8x8 square.
__device__ __forceinline__ void pixel_of_thread(int32_t tid, int32_t& py, int32_t& px) {
  const int32_t w = tid >> 6, l = tid & 63;
  px = ((w & 1) << 3) | (l & 7);
  py = ((w >> 1) << 3) | (l >> 3);
}

// Move codes recorded by the tracing pass (4 bits per step, 8 per word).
enum : uint32_t {
  kMoveStay = 0,
  kMoveXPlus = 1,
  kMoveXMinus = 2,
  kMoveYPlus = 3,
  kMoveYMinus = 4,
  kMoveStop = 7
};

}  // namespace rlic
