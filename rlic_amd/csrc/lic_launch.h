// Host-side launchers of the streamline kernels (defined in lic_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lic_device.h"

namespace rlic {

// wave_w_log2: log2 of the wave tile width (3: 8x8, 4: 16x4, 5: 32x2 pixels).
TileMap make_tile_map(int32_t ny, int32_t nx, int32_t band, int32_t wave_w_log2);

// Number of 16-byte move-code chunks (32 steps each) recorded per pixel
// slot for a kernel of length klen (forward chunks, then backward chunks).
int32_t code_chunks(int32_t klen);

// The field's guarded cell array (guarded_cells(ny, nx) cells) from its
// strided texture, u and v; with `kpad`, in the same launch, the padded
// weight copy of `kernel`: kpad (klen + 2 * kernel_pad() elements) =
// [zeros][kernel][zeros].
template <typename T>
hipError_t launch_pack(const T* tex, int64_t tsy, int64_t tsx, const T* u, int64_t usy,
                       int64_t usx, const T* v, int64_t vsy, int64_t vsx, const FieldGeom& g,
                       Cell<T>* cells, hipStream_t stream, const T* kernel = nullptr,
                       int32_t klen = 0, T* kpad = nullptr);

// Zeros on each side of the padded kernel-weight copy the streamline
// kernels read (they load taps in groups of 8 without bounds checks).
int32_t kernel_pad();

template <typename T>
hipError_t launch_repack_texture(const T* x, const FieldGeom& g, Cell<T>* cells,
                                 hipStream_t stream);

template <typename T>
hipError_t launch_copy2d(const T* src, int64_t sy, int64_t sx, int32_t ny, int32_t nx, T* dst,
                         hipStream_t stream);

// `kernel` points at the first weight of a padded copy (launch_pack).
// codes == nullptr: no recording.  Otherwise codes holds
// code_chunks(klen) * n_tiles * 256 chunks.
template <typename T>
hipError_t launch_trace(const Cell<T>* cells, const T* kernel, int32_t klen, const FieldGeom& g,
                        const TileMap& m, int uv_mode, int profile, T* out, uint4* codes,
                        hipStream_t stream);

// Paired trace (both directions of a pixel on two lanes) for fields too
// small to fill the GPU: trace_pair_fits tells whether it applies (odd
// klen, LDS for the backward lanes' values, buffer-addressable cells);
// launch_trace_pair computes what launch_trace does, bit for bit, with the
// same record, and returns hipErrorInvalidValue where it does not apply.
size_t trace_pair_lds_bytes(int32_t klen, size_t elem);
template <typename T>
bool trace_pair_fits(int32_t klen, int64_t ny, int64_t nx);
template <typename T>
hipError_t launch_trace_pair(const Cell<T>* cells, const T* kernel, int32_t klen,
                             const FieldGeom& g, const TileMap& m, int uv_mode, int profile,
                             T* out, uint4* codes, hipStream_t stream);

template <typename T>
hipError_t launch_replay(const T* x, const T* kernel, int32_t klen, const FieldGeom& g,
                         const TileMap& m, int profile, T* out, const uint4* codes,
                         hipStream_t stream);

// Synthetic inputs (lic_generate.hip): uniform noise in [0, 1) for n
// contiguous elements, and sums of plane waves on an ny x nx grid.
constexpr int32_t kMaxModes = 32;
struct ModeSet {
  int32_t n;                                  // modes in use (<= kMaxModes)
  double amp[kMaxModes], phase[kMaxModes];    // amplitude, phase (radians)
  double kx[kMaxModes], ky[kMaxModes];        // integer wavenumbers, as doubles
};
template <typename T>
hipError_t launch_fill_noise(T* out, int64_t n, uint64_t seed, hipStream_t stream);
template <typename T>
hipError_t launch_fill_modes(T* out, int32_t ny, int32_t nx, const ModeSet& modes,
                             hipStream_t stream);

// q[i] = a[i] / b[i] with the streamline kernels' float division (a check).
hipError_t launch_div_check(const float* a, const float* b, float* q, int64_t n,
                            hipStream_t stream);

// Adds the live lane-steps (stats[0]) and executed lane-steps (stats[1]) of
// a tracing pass with tile map m, read from its move record.
hipError_t launch_record_stats(const uint4* codes, int32_t klen, const FieldGeom& g,
                               const TileMap& m, unsigned long long* stats, hipStream_t stream);

// LDS-staged replay: tiles of 64x64, 32x64 or 32x32 pixels, the largest
// whose box (tile plus kmid halo) fits the device's per-block LDS.
struct LdsTile {
  int32_t h, w;
};
template <typename T>
LdsTile lds_replay_tile(int32_t klen);
template <typename T>
bool lds_replay_fits(int32_t klen);

template <typename T>
hipError_t launch_replay_lds(const T* x, const T* kernel, int32_t klen, const FieldGeom& g,
                             int profile, T* out, const uint4* codes, hipStream_t stream);

// The same with an explicit tile (64x64, 32x64 or 32x32); hipErrorInvalidValue
// when that tile's box does not fit.
template <typename T>
hipError_t launch_replay_lds_tile(const T* x, const T* kernel, int32_t klen, const FieldGeom& g,
                                  LdsTile tile, int profile, T* out, const uint4* codes,
                                  hipStream_t stream);

// Hybrid replay: 64x64 tiles whose box holds the first `halo` steps of each
// direction (32 or 64; at most 32 in double), later steps gathered from
// the image with buffer loads.  lds_hybrid_fits: the halo is one of those
// and below kmid, the box fits and the image (n elements) is
// buffer-addressable; otherwise the launch returns hipErrorInvalidValue.
template <typename T>
constexpr int32_t hybrid_max_halo() {
  return sizeof(T) == 4 ? 64 : 32;
}
template <typename T>
bool lds_hybrid_fits(int32_t klen, int64_t n, int32_t halo);
template <typename T>
hipError_t launch_replay_lds_hybrid(const T* x, const T* kernel, int32_t klen, const FieldGeom& g,
                                    int32_t halo, int profile, T* out, const uint4* codes,
                                    hipStream_t stream);

}  // namespace rlic
