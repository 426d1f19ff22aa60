// Host-side launchers of the streamline kernels (defined in lic_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lic_device.h"

namespace rlic {

TileMap make_tile_map(int32_t ny, int32_t nx, int32_t band);

// Number of 32-bit move-code words recorded per pixel for a kernel of
// length klen (forward and backward directions are word-aligned apart).
int32_t code_words(int32_t klen);

template <typename T>
hipError_t launch_pack(const T* tex, int64_t tsy, int64_t tsx, const T* u, int64_t usy,
                       int64_t usx, const T* v, int64_t vsy, int64_t vsx, int32_t ny,
                       int32_t nx, Cell<T>* cells, hipStream_t stream);

template <typename T>
hipError_t launch_repack_texture(const T* x, int32_t n, Cell<T>* cells, hipStream_t stream);

template <typename T>
hipError_t launch_copy2d(const T* src, int64_t sy, int64_t sx, int32_t ny, int32_t nx, T* dst,
                         hipStream_t stream);

// codes == nullptr: no recording.  Otherwise codes holds
// code_words(klen) * n_tiles * 256 words.
template <typename T>
hipError_t launch_trace(const Cell<T>* cells, const T* kernel, int32_t klen, const FieldGeom& g,
                        const TileMap& m, int uv_mode, int profile, T* out, uint32_t* codes,
                        hipStream_t stream);

template <typename T>
hipError_t launch_replay(const T* x, const T* kernel, int32_t klen, const FieldGeom& g,
                         const TileMap& m, int profile, T* out, const uint32_t* codes,
                         hipStream_t stream);

}  // namespace rlic
