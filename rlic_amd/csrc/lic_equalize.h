// Histogram equalization on the device (lic_equalize.hip): global, and
// adaptive by sliding window or by interpolated tiles.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace rlic {

// Bins whose per-block histogram is kept in LDS (larger bin counts use
// global atomics directly).
constexpr int32_t kLdsBins = 16384;

// Reduced range of the finite pixels: order-preserving integer keys of the
// smallest and largest value, and their count.
template <typename T>
struct EqState;
template <>
struct EqState<float> {
  uint32_t min_key, max_key;
  unsigned long long count;
};
template <>
struct EqState<double> {
  unsigned long long min_key, max_key;
  unsigned long long count;
};

// The binning derived from the range (see lic_equalize.hip).
struct EqRange {
  unsigned long long count;
  bool constant, halves;
  double lo, scale;
};

// Device scratch of one equalization: the range state and nbins counters.
template <typename T>
size_t equalize_scratch_bytes(int32_t nbins);

// Equalizes the ny x nx image (element strides sy, sx) into the C-order
// `out` (which may be the image itself when that is C-contiguous), all on
// `stream`; `scratch` holds equalize_scratch_bytes<T>(nbins) bytes.
template <typename T>
hipError_t launch_equalize(const T* img, int64_t sy, int64_t sx, int32_t ny, int32_t nx,
                           int32_t nbins, T* out, void* scratch, hipStream_t stream);

// Adaptive equalization (tile_h x tile_w windows or tiles; semantics in
// lic_equalize.hip): kEqSlidingTile counts each pixel's window,
// kEqTileInterpolation interpolates per-tile mappings between tile centres.
constexpr int kEqSlidingTile = 1;
constexpr int kEqTileInterpolation = 2;

template <typename T>
size_t equalize_tiles_scratch_bytes(int method, int32_t ny, int32_t nx, int32_t nbins,
                                    int32_t tile_h, int32_t tile_w);

// As launch_equalize, by `method`; `lds_limit`: the per-block LDS bytes
// available (a sliding window whose staged bins exceed it is counted from
// global memory).
template <typename T>
hipError_t launch_equalize_tiles(int method, const T* img, int64_t sy, int64_t sx, int32_t ny,
                                 int32_t nx, int32_t nbins, int32_t tile_h, int32_t tile_w, T* out,
                                 void* scratch, int32_t lds_limit, hipStream_t stream);

}  // namespace rlic
