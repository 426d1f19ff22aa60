// HIP kernels of the line integral convolution hot path (gfx950).
//
//   lic_pack    strided (texture, u, v) -> packed Cell array (one pass, HBM-bound)
//   lic_trace   one lane per output pixel: centre tap, forward streamline,
//               backward streamline (src/lib.rs:364-406, 304-362); optionally
//               records the cell-to-cell move of every step
//   lic_replay  later iterations: the vector field does not change between
//               iterations (src/lib.rs:416-420, 432-440), so every iteration
//               visits the same cells; replay walks the recorded moves and
//               only gathers the new texture, in the same summation order
//   lic_copy2d  strided 2-D copy (iterations == 0 and texture re-packing)
#include "lic_device.h"
#include "lic_launch.h"

namespace rlic {

template <typename T>
__global__ __launch_bounds__(256) void lic_pack(const T* __restrict__ tex, int64_t tsy,
                                                int64_t tsx, const T* __restrict__ u,
                                                int64_t usy, int64_t usx,
                                                const T* __restrict__ v, int64_t vsy,
                                                int64_t vsx, int32_t ny, int32_t nx,
                                                Cell<T>* __restrict__ cells) {
  const int32_t j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int32_t i = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= ny || j >= nx) return;
  Cell<T> c;
  c.u = u[i * usy + j * usx];
  c.v = v[i * vsy + j * vsx];
  c.x = tex[i * tsy + j * tsx];
  c.pad = T(0);
  cells[static_cast<int64_t>(i) * nx + j] = c;
}

// Overwrite the texture slot of the packed cells with a C-order array.
template <typename T>
__global__ __launch_bounds__(256) void lic_repack_texture(const T* __restrict__ x, int32_t n,
                                                          Cell<T>* __restrict__ cells) {
  const int32_t p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  cells[p].x = x[p];
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

// One streamline direction of one pixel (src/lib.rs:304-362), starting with
// the vector (pu, pv) of the start cell already loaded.  Returns the updated
// accumulator.  With REC, writes the move codes of this direction to
// codes[(w0 + s/8) * np + p].
template <typename T, bool FMA, bool BRANCHLESS, bool POL, bool FORWARD, bool REC>
__device__ __forceinline__ T trace_direction(T acc, T pu, T pv, int32_t i, int32_t j,
                                             const Cell<T>* __restrict__ cells,
                                             const T* __restrict__ kernel, int32_t kmid,
                                             int32_t nsteps, const FieldGeom& g,
                                             uint32_t* __restrict__ codes, int64_t np,
                                             int64_t p, int32_t w0) {
  T fx = T(0.5), fy = T(0.5);
  T last_u = T(0), last_v = T(0);
  uint32_t word = 0;
  for (int32_t s = 0; s < nsteps; ++s) {
    const int32_t k = FORWARD ? kmid + 1 + s : kmid - 1 - s;
    if (isnan_t(pu) || isnan_t(pv)) {  // src/lib.rs:336-338
      if constexpr (REC) {
        word |= kMoveStop << (4 * (s & 7));
        codes[(w0 + (s >> 3)) * np + p] = word;
      }
      break;
    }
    if constexpr (POL) {  // src/lib.rs:339-347, unfused dot product
      const T d1 = pu * last_u;
      const T d2 = pv * last_v;
      if (d1 + d2 < T(0)) {
        pu = -pu;
        pv = -pv;
      }
      last_u = pu;
      last_v = pv;
    }
    const T mu = FORWARD ? pu : -pu;  // src/lib.rs:348-351
    const T mv = FORWARD ? pv : -pv;
    uint32_t code = kMoveStay;
    if (!(mu == T(0) && mv == T(0))) {  // src/lib.rs:242-244
      const T tx = time_to_next_pixel<T, FMA, BRANCHLESS>(mu, fx);
      const T ty = time_to_next_pixel<T, FMA, BRANCHLESS>(mv, fy);
      if (tx < ty) {  // src/lib.rs:249-258, update_state :208-233
        if (mu >= T(0)) {
          j += 1;
          fx = T(0);
          code = kMoveXPlus;
        } else {
          j -= 1;
          fx = T(1);
          code = kMoveXMinus;
        }
        fy = macc<FMA>(tx, mv, fy);
      } else {  // src/lib.rs:259-269
        if (mv >= T(0)) {
          i += 1;
          fy = T(0);
          code = kMoveYPlus;
        } else {
          i -= 1;
          fy = T(1);
          code = kMoveYMinus;
        }
        fx = macc<FMA>(ty, mu, fx);
      }
      j = apply_bound(j, g.nx, g.x_left, g.x_right);  // src/lib.rs:112-115
      i = apply_bound(i, g.ny, g.y_left, g.y_right);
    }
    const CellVal<T> c = load_cell(cells, i * g.nx + j);
    acc = macc<FMA>(kernel[k], c.x, acc);  // src/lib.rs:353-360
    pu = c.u;
    pv = c.v;
    if constexpr (REC) {
      word |= code << (4 * (s & 7));
      if ((s & 7) == 7 || s == nsteps - 1) {
        codes[(w0 + (s >> 3)) * np + p] = word;
        word = 0;
      }
    }
  }
  return acc;
}

template <typename T, bool FMA, bool BRANCHLESS, bool POL, bool REC>
__global__ __launch_bounds__(256) void lic_trace(const Cell<T>* __restrict__ cells,
                                                 const T* __restrict__ kernel, int32_t klen,
                                                 FieldGeom g, TileMap m, T* __restrict__ out,
                                                 uint32_t* __restrict__ codes) {
  int32_t ty, tx, py, px;
  tile_of_block(blockIdx.x, m, ty, tx);
  pixel_of_thread(threadIdx.x, py, px);
  const int32_t i0 = ty * kTile + py, j0 = tx * kTile + px;
  if (i0 >= g.ny || j0 >= g.nx) return;
  const int32_t kmid = klen / 2;
  const int32_t n_fwd = klen - 1 - kmid, n_bwd = kmid;
  const int64_t np = static_cast<int64_t>(m.n_tiles) * kBlock;
  const int64_t p = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  const int32_t idx0 = i0 * g.nx + j0;
  const CellVal<T> c0 = load_cell(cells, idx0);
  T acc = macc<FMA>(kernel[kmid], c0.x, T(0));  // src/lib.rs:375-383
  acc = trace_direction<T, FMA, BRANCHLESS, POL, true, REC>(
      acc, c0.u, c0.v, i0, j0, cells, kernel, kmid, n_fwd, g, codes, np, p, 0);
  acc = trace_direction<T, FMA, BRANCHLESS, POL, false, REC>(
      acc, c0.u, c0.v, i0, j0, cells, kernel, kmid, n_bwd, g, codes, np, p,
      (n_fwd + 7) >> 3);
  out[idx0] = acc;
}

template <typename T, bool FMA, bool FORWARD>
__device__ __forceinline__ T replay_direction(T acc, int32_t i, int32_t j,
                                              const T* __restrict__ x,
                                              const T* __restrict__ kernel, int32_t kmid,
                                              int32_t nsteps, const FieldGeom& g,
                                              const uint32_t* __restrict__ codes, int64_t np,
                                              int64_t p, int32_t w0) {
  uint32_t word = 0;
  for (int32_t s = 0; s < nsteps; ++s) {
    const int32_t k = FORWARD ? kmid + 1 + s : kmid - 1 - s;
    if ((s & 7) == 0) word = codes[(w0 + (s >> 3)) * np + p];
    const uint32_t code = (word >> (4 * (s & 7))) & 15u;
    if (code == kMoveStop) break;
    j += (code == kMoveXPlus) ? 1 : ((code == kMoveXMinus) ? -1 : 0);
    i += (code == kMoveYPlus) ? 1 : ((code == kMoveYMinus) ? -1 : 0);
    j = apply_bound(j, g.nx, g.x_left, g.x_right);
    i = apply_bound(i, g.ny, g.y_left, g.y_right);
    acc = macc<FMA>(kernel[k], x[i * g.nx + j], acc);
  }
  return acc;
}

template <typename T, bool FMA>
__global__ __launch_bounds__(256) void lic_replay(const T* __restrict__ x,
                                                  const T* __restrict__ kernel, int32_t klen,
                                                  FieldGeom g, TileMap m, T* __restrict__ out,
                                                  const uint32_t* __restrict__ codes) {
  int32_t ty, tx, py, px;
  tile_of_block(blockIdx.x, m, ty, tx);
  pixel_of_thread(threadIdx.x, py, px);
  const int32_t i0 = ty * kTile + py, j0 = tx * kTile + px;
  if (i0 >= g.ny || j0 >= g.nx) return;
  const int32_t kmid = klen / 2;
  const int32_t n_fwd = klen - 1 - kmid, n_bwd = kmid;
  const int64_t np = static_cast<int64_t>(m.n_tiles) * kBlock;
  const int64_t p = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  const int32_t idx0 = i0 * g.nx + j0;
  T acc = macc<FMA>(kernel[kmid], x[idx0], T(0));
  acc = replay_direction<T, FMA, true>(acc, i0, j0, x, kernel, kmid, n_fwd, g, codes, np, p, 0);
  acc = replay_direction<T, FMA, false>(acc, i0, j0, x, kernel, kmid, n_bwd, g, codes, np, p,
                                        (n_fwd + 7) >> 3);
  out[idx0] = acc;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

TileMap make_tile_map(int32_t ny, int32_t nx, int32_t band) {
  TileMap m;
  m.tiles_x = (nx + kTile - 1) / kTile;
  m.tiles_y = (ny + kTile - 1) / kTile;
  m.n_tiles = m.tiles_x * m.tiles_y;
  m.per_xcd = m.n_tiles / kXcds;
  m.rem_xcd = m.n_tiles % kXcds;
  m.band = band < 1 ? 1 : band;
  return m;
}

int32_t code_words(int32_t klen) {
  const int32_t kmid = klen / 2;
  const int32_t n_fwd = klen - 1 - kmid, n_bwd = kmid;
  return ((n_fwd + 7) >> 3) + ((n_bwd + 7) >> 3);
}

template <typename T>
hipError_t launch_pack(const T* tex, int64_t tsy, int64_t tsx, const T* u, int64_t usy,
                       int64_t usx, const T* v, int64_t vsy, int64_t vsx, int32_t ny,
                       int32_t nx, Cell<T>* cells, hipStream_t stream) {
  dim3 grid((nx + 63) / 64, (ny + 3) / 4);
  hipLaunchKernelGGL(lic_pack<T>, grid, dim3(256), 0, stream, tex, tsy, tsx, u, usy, usx, v,
                     vsy, vsx, ny, nx, cells);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_repack_texture(const T* x, int32_t n, Cell<T>* cells, hipStream_t stream) {
  hipLaunchKernelGGL(lic_repack_texture<T>, dim3((n + 255) / 256), dim3(256), 0, stream, x, n,
                     cells);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_copy2d(const T* src, int64_t sy, int64_t sx, int32_t ny, int32_t nx, T* dst,
                         hipStream_t stream) {
  dim3 grid((nx + 63) / 64, (ny + 3) / 4);
  hipLaunchKernelGGL(lic_copy2d<T>, grid, dim3(256), 0, stream, src, sy, sx, ny, nx, dst);
  return hipGetLastError();
}

template <typename T, bool FMA, bool BL, bool POL, bool REC>
static hipError_t trace_one(const Cell<T>* cells, const T* kernel, int32_t klen,
                            const FieldGeom& g, const TileMap& m, T* out, uint32_t* codes,
                            hipStream_t stream) {
  hipLaunchKernelGGL((lic_trace<T, FMA, BL, POL, REC>), dim3(m.n_tiles), dim3(kBlock), 0,
                     stream, cells, kernel, klen, g, m, out, codes);
  return hipGetLastError();
}

template <typename T, bool FMA, bool BL>
static hipError_t trace_mode(const Cell<T>* cells, const T* kernel, int32_t klen,
                             const FieldGeom& g, const TileMap& m, int uv_mode, T* out,
                             uint32_t* codes, hipStream_t stream) {
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
                        const TileMap& m, int uv_mode, int profile, T* out, uint32_t* codes,
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

template <typename T>
hipError_t launch_replay(const T* x, const T* kernel, int32_t klen, const FieldGeom& g,
                         const TileMap& m, int profile, T* out, const uint32_t* codes,
                         hipStream_t stream) {
  if (profile & 1) {
    hipLaunchKernelGGL((lic_replay<T, true>), dim3(m.n_tiles), dim3(kBlock), 0, stream, x,
                       kernel, klen, g, m, out, codes);
  } else {
    hipLaunchKernelGGL((lic_replay<T, false>), dim3(m.n_tiles), dim3(kBlock), 0, stream, x,
                       kernel, klen, g, m, out, codes);
  }
  return hipGetLastError();
}

#define RLIC_INSTANTIATE(T)                                                                     \
  template hipError_t launch_pack<T>(const T*, int64_t, int64_t, const T*, int64_t, int64_t,   \
                                     const T*, int64_t, int64_t, int32_t, int32_t, Cell<T>*,   \
                                     hipStream_t);                                             \
  template hipError_t launch_repack_texture<T>(const T*, int32_t, Cell<T>*, hipStream_t);      \
  template hipError_t launch_copy2d<T>(const T*, int64_t, int64_t, int32_t, int32_t, T*,       \
                                       hipStream_t);                                           \
  template hipError_t launch_trace<T>(const Cell<T>*, const T*, int32_t, const FieldGeom&,     \
                                      const TileMap&, int, int, T*, uint32_t*, hipStream_t);   \
  template hipError_t launch_replay<T>(const T*, const T*, int32_t, const FieldGeom&,          \
                                       const TileMap&, int, T*, const uint32_t*, hipStream_t);

RLIC_INSTANTIATE(float)
RLIC_INSTANTIATE(double)

}  // namespace rlic
