// Global histogram equalization of an image on the device: the post-step a
// LIC image usually gets, applied where the LIC output already lies (HBM),
// so the image never travels to the host in between.
//
// The reference declares `equalize_histogram_f32/_f64(image, nbins)`
// (src/rlic/_core.pyi:30-37) without implementing it anywhere; these
// kernels implement the textbook mapping, out(x) = cdf(bin(x)) / N:
//   * finite pixels take part; NaN and +-inf pixels are left out of the
//     range and the histogram and are copied unchanged;
//   * vmin, vmax = smallest and largest finite value, N = finite pixels;
//   * bin(x) = min(floor((x - vmin) * (nbins / (vmax - vmin))), nbins - 1)
//     in IEEE double (float pixels widened exactly); when vmax - vmin
//     overflows, with halves: (x/2 - vmin/2) * (nbins / (vmax/2 - vmin/2));
//     when vmax == vmin every finite pixel is in bin 0;
//   * out = cdf[bin] / N in double, rounded once to the image type.
//
// Four stream-ordered launches, no host round trip: the finite range (one
// read of the image; block min/max/count, then one atomic per block on
// order-preserving integer keys), the histogram (a second read; per-block
// LDS histograms merged with global atomics), an inclusive scan of the
// counts (one block), and the map (a third read and the write).  All of
// it is integer and HBM traffic: 3 reads and 1 write of the image.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>

#include "lic_equalize.h"

namespace rlic {

namespace {

// Unsigned keys whose order is the float order (for atomic min / max).
__device__ __forceinline__ uint32_t order_key(float x) {
  const uint32_t b = __builtin_bit_cast(uint32_t, x);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ unsigned long long order_key(double x) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ float from_key(uint32_t k) {
  return __builtin_bit_cast(float, (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}
__device__ __forceinline__ double from_key(unsigned long long k) {
  return __builtin_bit_cast(double, (k & 0x8000000000000000ull) ? (k & 0x7FFFFFFFFFFFFFFFull)
                                                                 : ~k);
}

template <typename T>
struct KeyOf;
template <>
struct KeyOf<float> {
  typedef uint32_t type;
  static constexpr uint32_t kMax = 0xFFFFFFFFu;
};
template <>
struct KeyOf<double> {
  typedef unsigned long long type;
  static constexpr unsigned long long kMax = 0xFFFFFFFFFFFFFFFFull;
};

constexpr int kThreads = 256;

template <typename T>
__device__ __forceinline__ T load_px(const T* __restrict__ img, int64_t p, int32_t nx, int64_t sy,
                                     int64_t sx) {
  const int64_t i = p / nx, j = p - i * nx;
  return img[i * sy + j * sx];
}

// The bin of a finite pixel (see the file comment); vmin < vmax.
__device__ __forceinline__ int32_t eq_bin(double x, const EqRange& r, int32_t nbins) {
  const double t = r.halves ? (x * 0.5 - r.lo) * r.scale : (x - r.lo) * r.scale;
  const double f = floor(t);
  return f >= static_cast<double>(nbins - 1) ? nbins - 1 : static_cast<int32_t>(f);
}

// The range and scale of the binning, from the reduced keys (every thread
// computes the same values).
template <typename T>
__device__ __forceinline__ EqRange read_range(const EqState<T>* st, int32_t nbins) {
  EqRange r;
  r.count = st->count;
  r.constant = true;
  r.halves = false;
  r.lo = 0.0;
  r.scale = 0.0;
  if (r.count == 0) return r;
  const double vmin = static_cast<double>(from_key(st->min_key));
  const double vmax = static_cast<double>(from_key(st->max_key));
  if (!(vmax > vmin)) return r;
  r.constant = false;
  const double d = vmax - vmin;
  if (d <= 1.7976931348623157e308) {
    r.lo = vmin;
    r.scale = static_cast<double>(nbins) / d;
  } else {  // the difference overflows: both halved
    r.halves = true;
    r.lo = vmin * 0.5;
    r.scale = static_cast<double>(nbins) / (vmax * 0.5 - vmin * 0.5);
  }
  return r;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void lic_eq_init(EqState<T>* st, uint32_t* hist,
                                                       int32_t nbins) {
  const int32_t t = blockIdx.x * kThreads + threadIdx.x;
  if (t == 0) {
    st->min_key = KeyOf<T>::kMax;
    st->max_key = 0;
    st->count = 0;
  }
  for (int32_t b = t; b < nbins; b += gridDim.x * kThreads) hist[b] = 0u;
}

// Smallest and largest finite value and the count of finite pixels.
template <typename T>
__global__ __launch_bounds__(kThreads) void lic_eq_range(const T* __restrict__ img, int64_t n,
                                                        int32_t nx, int64_t sy, int64_t sx,
                                                        EqState<T>* st) {
  typedef typename KeyOf<T>::type Key;
  Key lo = KeyOf<T>::kMax, hi = 0;
  unsigned long long cnt = 0;
  for (int64_t p = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; p < n;
       p += static_cast<int64_t>(gridDim.x) * kThreads) {
    const T x = load_px(img, p, nx, sy, sx);
    if (__builtin_isfinite(x)) {
      const Key k = order_key(x);
      lo = k < lo ? k : lo;
      hi = k > hi ? k : hi;
      ++cnt;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const Key olo = __shfl_xor(lo, off), ohi = __shfl_xor(hi, off);
    lo = olo < lo ? olo : lo;
    hi = ohi > hi ? ohi : hi;
    cnt += __shfl_xor(cnt, off);
  }
  __shared__ Key s_lo[kThreads / 64], s_hi[kThreads / 64];
  __shared__ unsigned long long s_cnt[kThreads / 64];
  const int32_t w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_lo[w] = lo;
    s_hi[w] = hi;
    s_cnt[w] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kThreads / 64; ++k) {
      lo = s_lo[k] < lo ? s_lo[k] : lo;
      hi = s_hi[k] > hi ? s_hi[k] : hi;
      cnt += s_cnt[k];
    }
    if (cnt != 0) {
      atomicMin(&st->min_key, lo);
      atomicMax(&st->max_key, hi);
      atomicAdd(&st->count, cnt);
    }
  }
}

// Histogram of the finite pixels' bins: per-block LDS histograms (nbins up
// to kLdsBins) merged into the global one, else global atomics.
template <typename T, bool LDS>
__global__ __launch_bounds__(kThreads) void lic_eq_hist(const T* __restrict__ img, int64_t n,
                                                       int32_t nx, int64_t sy, int64_t sx,
                                                       int32_t nbins, const EqState<T>* st,
                                                       uint32_t* __restrict__ hist) {
  __shared__ uint32_t local[LDS ? kLdsBins : 1];
  const EqRange r = read_range(st, nbins);
  if (LDS) {
    for (int32_t b = threadIdx.x; b < nbins; b += kThreads) local[b] = 0u;
    __syncthreads();
  }
  for (int64_t p = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; p < n;
       p += static_cast<int64_t>(gridDim.x) * kThreads) {
    const T x = load_px(img, p, nx, sy, sx);
    if (!__builtin_isfinite(x)) continue;
    const int32_t b = r.constant ? 0 : eq_bin(static_cast<double>(x), r, nbins);
    if (LDS) {
      atomicAdd(&local[b], 1u);
    } else {
      atomicAdd(&hist[b], 1u);
    }
  }
  if (LDS) {
    __syncthreads();
    for (int32_t b = threadIdx.x; b < nbins; b += kThreads)
      if (local[b] != 0u) atomicAdd(&hist[b], local[b]);
  }
}

// Inclusive scan of the counts, in place: block k scans the k-th array of
// nbins counters (one block for the global histogram, one per tile).
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void lic_eq_scan(uint32_t* hist_all, int32_t nbins) {
  uint32_t* hist = hist_all + static_cast<int64_t>(blockIdx.x) * nbins;
  __shared__ uint32_t part[kScanThreads / 64];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0u;
  __syncthreads();
  const int32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int32_t base = 0; base < nbins; base += kScanThreads) {
    const int32_t b = base + threadIdx.x;
    uint32_t v = b < nbins ? hist[b] : 0u;
    for (int off = 1; off < 64; off <<= 1) {  // inclusive scan within the wave
      const uint32_t o = __shfl_up(v, off);
      if (lane >= off) v += o;
    }
    if (lane == 63) part[w] = v;
    __syncthreads();
    uint32_t before = carry;
    for (int k = 0; k < w; ++k) before += part[k];
    if (b < nbins) hist[b] = v + before;
    __syncthreads();
    if (threadIdx.x == kScanThreads - 1) carry = v + before;
    __syncthreads();
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void lic_eq_map(const T* __restrict__ img, int64_t n,
                                                      int32_t nx, int64_t sy, int64_t sx,
                                                      int32_t nbins, const EqState<T>* st,
                                                      const uint32_t* __restrict__ cdf,
                                                      T* __restrict__ out) {
  const EqRange r = read_range(st, nbins);
  for (int64_t p = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; p < n;
       p += static_cast<int64_t>(gridDim.x) * kThreads) {
    const T x = load_px(img, p, nx, sy, sx);
    T y = x;
    if (__builtin_isfinite(x)) {
      const int32_t b = r.constant ? 0 : eq_bin(static_cast<double>(x), r, nbins);
      y = static_cast<T>(static_cast<double>(cdf[b]) / static_cast<double>(r.count));
    }
    out[p] = y;
  }
}

// ---------------------------------------------------------------------------
// Adaptive equalization (the reference's sliding-tile and tile-interpolation
// stubs, src/rlic/_core.pyi:38-57; semantics in oracle/equalize.py): the
// global bins above, counted per window or per tile.
// ---------------------------------------------------------------------------

// Bin of every pixel (-1 where not finite), C order.
template <typename T>
__global__ __launch_bounds__(kThreads) void lic_eq_bins(const T* __restrict__ img, int64_t n,
                                                       int32_t nx, int64_t sy, int64_t sx,
                                                       int32_t nbins, const EqState<T>* st,
                                                       int32_t* __restrict__ bins) {
  const EqRange r = read_range(st, nbins);
  for (int64_t p = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; p < n;
       p += static_cast<int64_t>(gridDim.x) * kThreads) {
    const T x = load_px(img, p, nx, sy, sx);
    bins[p] = __builtin_isfinite(x) ? (r.constant ? 0 : eq_bin(static_cast<double>(x), r, nbins))
                                    : -1;
  }
}

// Sliding tile: out = (finite window pixels with a bin <= the pixel's) /
// (finite window pixels).  A block covers kSlide x kSlide pixels; with LDS
// it first stages the bins of its pixels' windows (the block plus th - 1
// rows and tw - 1 columns, -1 outside the image) as dynamic LDS.  A bin
// compares as unsigned, so -1 (not finite) is never "<=" a pixel's bin.
constexpr int32_t kSlide = 16;

template <typename T, bool LDS>
__global__ __launch_bounds__(kSlide * kSlide) void lic_eq_sliding(
    const T* __restrict__ img, int32_t ny, int32_t nx, int64_t sy, int64_t sx,
    const int32_t* __restrict__ bins, int32_t th, int32_t tw, T* __restrict__ out) {
  extern __shared__ int32_t box[];
  const int32_t up = (th - 1) / 2, left = (tw - 1) / 2;
  const int32_t i0 = blockIdx.y * kSlide, j0 = blockIdx.x * kSlide;
  // the box's sides (LDS launches only: their box fits the LDS)
  const int32_t bw = LDS ? kSlide + tw - 1 : 0;
  if (LDS) {
    const int32_t bh = kSlide + th - 1;
    for (int32_t k = threadIdx.x; k < bh * bw; k += kSlide * kSlide) {
      const int32_t r = k / bw, c = k - r * bw;
      const int32_t i = i0 - up + r, j = j0 - left + c;
      box[k] = (i >= 0 && i < ny && j >= 0 && j < nx) ? bins[static_cast<int64_t>(i) * nx + j]
                                                        : -1;
    }
    __syncthreads();
  }
  const int32_t ti = threadIdx.x / kSlide, tj = threadIdx.x % kSlide;
  const int32_t i = i0 + ti, j = j0 + tj;
  if (i >= ny || j >= nx) return;
  const int64_t p = static_cast<int64_t>(i) * nx + j;
  const T x = img[i * sy + j * sx];
  const int32_t bp = bins[p];
  if (bp < 0) {
    out[p] = x;
    return;
  }
  uint32_t below = 0, count = 0;
  if (LDS) {
    for (int32_t r = 0; r < th; ++r) {
      const int32_t* row = box + (ti + r) * bw + tj;
      for (int32_t c = 0; c < tw; ++c) {
        const int32_t q = row[c];
        count += q >= 0 ? 1u : 0u;
        below += static_cast<uint32_t>(q) <= static_cast<uint32_t>(bp) ? 1u : 0u;
      }
    }
  } else {
    // (64-bit: i - up + th can pass 2^31 for windows near 2^31 rows)
    const int32_t r0 = max(0, i - up);
    const int32_t r1 = static_cast<int32_t>(min<int64_t>(ny, int64_t{i} - up + th));
    const int32_t c0 = max(0, j - left);
    const int32_t c1 = static_cast<int32_t>(min<int64_t>(nx, int64_t{j} - left + tw));
    for (int32_t r = r0; r < r1; ++r) {
      const int32_t* row = bins + static_cast<int64_t>(r) * nx;
      for (int32_t c = c0; c < c1; ++c) {
        const int32_t q = row[c];
        count += q >= 0 ? 1u : 0u;
        below += static_cast<uint32_t>(q) <= static_cast<uint32_t>(bp) ? 1u : 0u;
      }
    }
  }
  out[p] = static_cast<T>(static_cast<double>(below) / static_cast<double>(count));
}

// Tile interpolation.  Tile t = ty * tiles_x + tx covers rows
// [ty * th, min((ty + 1) * th, ny)) and columns alike; its counts go to
// hist[t * nbins ..].  With LDS a block counts its tile in LDS and writes
// every bin; without, the counters were cleared first and it adds with
// global atomics.
struct TileGrid {
  int32_t th, tw, tiles_y, tiles_x;
};

template <typename T, bool LDS>
__global__ __launch_bounds__(kThreads) void lic_eq_tile_hist(
    const T* __restrict__ img, int32_t ny, int32_t nx, int64_t sy, int64_t sx, int32_t nbins,
    const EqState<T>* st, TileGrid tg, uint32_t* __restrict__ hist_all) {
  __shared__ uint32_t local[LDS ? kLdsBins : 1];
  const EqRange r = read_range(st, nbins);
  const int32_t t = blockIdx.x;
  const int32_t ty = t / tg.tiles_x, tx = t - ty * tg.tiles_x;
  uint32_t* hist = hist_all + static_cast<int64_t>(t) * nbins;
  if (LDS) {
    for (int32_t b = threadIdx.x; b < nbins; b += kThreads) local[b] = 0u;
    __syncthreads();
  }
  const int32_t i0 = ty * tg.th, j0 = tx * tg.tw;
  const int32_t hh = min(tg.th, ny - i0), ww = min(tg.tw, nx - j0);
  const int64_t cells = static_cast<int64_t>(hh) * ww;
  for (int64_t k = threadIdx.x; k < cells; k += kThreads) {
    const int64_t a = k / ww;
    const int64_t i = i0 + a, j = j0 + (k - a * ww);
    const T x = img[i * sy + j * sx];
    if (!__builtin_isfinite(x)) continue;
    const int32_t b = r.constant ? 0 : eq_bin(static_cast<double>(x), r, nbins);
    if (LDS) {
      atomicAdd(&local[b], 1u);
    } else {
      atomicAdd(&hist[b], 1u);
    }
  }
  if (LDS) {
    __syncthreads();
    for (int32_t b = threadIdx.x; b < nbins; b += kThreads) hist[b] = local[b];
  }
}

__global__ __launch_bounds__(kThreads) void lic_eq_zero(uint32_t* __restrict__ p, int64_t n) {
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; k < n;
       k += static_cast<int64_t>(gridDim.x) * kThreads)
    p[k] = 0u;
}

// Tiles k, k + 1 and the weight of k + 1 for index i along an axis of n
// pixels cut into tiles of t (centre of tile k: k t + (extent - 1) / 2).
__device__ __forceinline__ double tile_centre(int32_t k, int32_t t, int32_t n) {
  const int32_t extent = min(t, n - k * t);
  return static_cast<double>(static_cast<int64_t>(k) * t) + static_cast<double>(extent - 1) * 0.5;
}

__device__ __forceinline__ void axis_weights(int32_t i, int32_t n, int32_t t, int32_t tiles,
                                             int32_t& a0, int32_t& a1, double& w) {
  const double x = static_cast<double>(i);
  a0 = a1 = 0;
  w = 0.0;
  if (tiles == 1 || x <= tile_centre(0, t, n)) return;
  if (x >= tile_centre(tiles - 1, t, n)) {
    a0 = a1 = tiles - 1;
    return;
  }
  // the last k with centre(k) <= x, in [0, tiles - 2]: a first guess from
  // full tiles, then corrected (the last tile may be partial)
  int32_t k = static_cast<int32_t>((x - static_cast<double>(t - 1) * 0.5) / t);
  k = k < 0 ? 0 : (k > tiles - 2 ? tiles - 2 : k);
  while (k < tiles - 2 && tile_centre(k + 1, t, n) <= x) ++k;
  while (k > 0 && tile_centre(k, t, n) > x) --k;
  a0 = k;
  a1 = k + 1;
  const double c0 = tile_centre(k, t, n);
  w = (x - c0) / (tile_centre(k + 1, t, n) - c0);
}

__device__ __forceinline__ double lerp_d(double a, double b, double w) { return a + w * (b - a); }

template <typename T>
__global__ __launch_bounds__(kThreads) void lic_eq_tile_map(
    const T* __restrict__ img, int32_t ny, int32_t nx, int64_t sy, int64_t sx, int32_t nbins,
    const EqState<T>* st, const uint32_t* __restrict__ gcdf, const uint32_t* __restrict__ tcdf,
    TileGrid tg, T* __restrict__ out) {
  const EqRange r = read_range(st, nbins);
  const int64_t n = static_cast<int64_t>(ny) * nx;
  for (int64_t p = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; p < n;
       p += static_cast<int64_t>(gridDim.x) * kThreads) {
    const int32_t i = static_cast<int32_t>(p / nx), j = static_cast<int32_t>(p - int64_t{i} * nx);
    const T x = img[i * sy + j * sx];
    T y = x;
    if (__builtin_isfinite(x)) {
      const int32_t b = r.constant ? 0 : eq_bin(static_cast<double>(x), r, nbins);
      int32_t y0, y1, x0, x1;
      double wy, wx;
      axis_weights(i, ny, tg.th, tg.tiles_y, y0, y1, wy);
      axis_weights(j, nx, tg.tw, tg.tiles_x, x0, x1, wx);
      auto map = [&](int32_t ty, int32_t tx) {
        const uint32_t* c = tcdf + (static_cast<int64_t>(ty) * tg.tiles_x + tx) * nbins;
        const uint32_t total = c[nbins - 1];
        return total != 0u ? static_cast<double>(c[b]) / static_cast<double>(total)
                           : static_cast<double>(gcdf[b]) / static_cast<double>(r.count);
      };
      const double top = lerp_d(map(y0, x0), map(y0, x1), wx);
      const double bot = lerp_d(map(y1, x0), map(y1, x1), wx);
      y = static_cast<T>(lerp_d(top, bot, wy));
    }
    out[p] = y;
  }
}

int grid_for(int64_t n, int32_t per_cu) {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    (void)hipGetLastError();
    cus = 256;
  }
  const int64_t want = (n + kThreads * 8 - 1) / (kThreads * 8);
  const int64_t cap = static_cast<int64_t>(cus) * per_cu;
  return static_cast<int>(want < 1 ? 1 : (want > cap ? cap : want));
}

}  // namespace

template <typename T>
size_t equalize_scratch_bytes(int32_t nbins) {
  return ((sizeof(EqState<T>) + 255) & ~size_t{255}) + sizeof(uint32_t) * static_cast<size_t>(nbins);
}

template <typename T>
hipError_t launch_equalize(const T* img, int64_t sy, int64_t sx, int32_t ny, int32_t nx,
                           int32_t nbins, T* out, void* scratch, hipStream_t stream) {
  const int64_t n = static_cast<int64_t>(ny) * nx;
  auto* st = static_cast<EqState<T>*>(scratch);
  auto* hist = reinterpret_cast<uint32_t*>(static_cast<char*>(scratch) +
                                           ((sizeof(EqState<T>) + 255) & ~size_t{255}));
  const int init_blocks = static_cast<int>((nbins + kThreads - 1) / kThreads > 1024
                                               ? 1024
                                               : (nbins + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(lic_eq_init<T>, dim3(init_blocks < 1 ? 1 : init_blocks), dim3(kThreads), 0,
                     stream, st, hist, nbins);
  const int g = grid_for(n, 8);
  hipLaunchKernelGGL(lic_eq_range<T>, dim3(g), dim3(kThreads), 0, stream, img, n, nx, sy, sx, st);
  // LDS histograms cost nbins of initialisation and merging per block:
  // fewer, longer blocks when nbins is large
  const int gh = grid_for(n, nbins <= 1024 ? 8 : 2);
  if (nbins <= kLdsBins) {
    hipLaunchKernelGGL((lic_eq_hist<T, true>), dim3(gh), dim3(kThreads), 0, stream, img, n, nx,
                       sy, sx, nbins, st, hist);
  } else {
    hipLaunchKernelGGL((lic_eq_hist<T, false>), dim3(gh), dim3(kThreads), 0, stream, img, n, nx,
                       sy, sx, nbins, st, hist);
  }
  hipLaunchKernelGGL(lic_eq_scan, dim3(1), dim3(kScanThreads), 0, stream, hist, nbins);
  hipLaunchKernelGGL(lic_eq_map<T>, dim3(g), dim3(kThreads), 0, stream, img, n, nx, sy, sx, nbins,
                     st, hist, out);
  return hipGetLastError();
}

int32_t tiles_along(int32_t n, int32_t t) { return (n + t - 1) / t; }

template <typename T>
size_t equalize_tiles_scratch_bytes(int method, int32_t ny, int32_t nx, int32_t nbins,
                                    int32_t th, int32_t tw) {
  size_t b = equalize_scratch_bytes<T>(nbins);
  b = (b + 255) & ~size_t{255};
  if (method == kEqSlidingTile)
    return b + sizeof(int32_t) * static_cast<size_t>(ny) * static_cast<size_t>(nx);
  const size_t tiles = static_cast<size_t>(tiles_along(ny, th)) * tiles_along(nx, tw);
  return b + sizeof(uint32_t) * tiles * static_cast<size_t>(nbins);
}

template <typename T>
hipError_t launch_equalize_tiles(int method, const T* img, int64_t sy, int64_t sx, int32_t ny,
                                 int32_t nx, int32_t nbins, int32_t th, int32_t tw, T* out,
                                 void* scratch, int32_t lds_limit, hipStream_t stream) {
  const int64_t n = static_cast<int64_t>(ny) * nx;
  auto* st = static_cast<EqState<T>*>(scratch);
  auto* hist = reinterpret_cast<uint32_t*>(static_cast<char*>(scratch) +
                                           ((sizeof(EqState<T>) + 255) & ~size_t{255}));
  char* extra = static_cast<char*>(scratch) +
                ((equalize_scratch_bytes<T>(nbins) + 255) & ~size_t{255});
  const int init_blocks = static_cast<int>((nbins + kThreads - 1) / kThreads > 1024
                                               ? 1024
                                               : (nbins + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(lic_eq_init<T>, dim3(init_blocks < 1 ? 1 : init_blocks), dim3(kThreads), 0,
                     stream, st, hist, nbins);
  const int g = grid_for(n, 8);
  hipLaunchKernelGGL(lic_eq_range<T>, dim3(g), dim3(kThreads), 0, stream, img, n, nx, sy, sx, st);
  if (method == kEqSlidingTile) {
    // a window taller than 2 ny - 1 rows covers the image from every pixel,
    // as one of exactly 2 ny - 1 does (likewise for columns)
    th = static_cast<int32_t>(std::min<int64_t>(th, 2 * int64_t{ny} - 1));
    tw = static_cast<int32_t>(std::min<int64_t>(tw, 2 * int64_t{nx} - 1));
    auto* bins = reinterpret_cast<int32_t*>(extra);
    hipLaunchKernelGGL(lic_eq_bins<T>, dim3(g), dim3(kThreads), 0, stream, img, n, nx, sy, sx,
                       nbins, st, bins);
    const dim3 grid(static_cast<unsigned>((nx + kSlide - 1) / kSlide),
                    static_cast<unsigned>((ny + kSlide - 1) / kSlide));
    const int64_t box = (int64_t{kSlide} + th - 1) * (int64_t{kSlide} + tw - 1) *
                        static_cast<int64_t>(sizeof(int32_t));
    if (box <= lds_limit) {
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&lic_eq_sliding<T, true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(box));
      (void)hipGetLastError();  // advisory on this target
      hipLaunchKernelGGL((lic_eq_sliding<T, true>), grid, dim3(kSlide * kSlide),
                         static_cast<unsigned>(box), stream, img, ny, nx, sy, sx, bins, th, tw,
                         out);
    } else {
      hipLaunchKernelGGL((lic_eq_sliding<T, false>), grid, dim3(kSlide * kSlide), 0, stream, img,
                         ny, nx, sy, sx, bins, th, tw, out);
    }
    return hipGetLastError();
  }
  // tile interpolation: the global histogram (for tiles without a finite
  // pixel), the tiles' histograms, their cumulative sums, the map
  const TileGrid tg{std::min(th, ny), std::min(tw, nx), tiles_along(ny, std::min(th, ny)),
                    tiles_along(nx, std::min(tw, nx))};
  const int32_t tiles = tg.tiles_y * tg.tiles_x;
  auto* tcdf = reinterpret_cast<uint32_t*>(extra);
  const int gh = grid_for(n, nbins <= 1024 ? 8 : 2);
  if (nbins <= kLdsBins) {
    hipLaunchKernelGGL((lic_eq_hist<T, true>), dim3(gh), dim3(kThreads), 0, stream, img, n, nx,
                       sy, sx, nbins, st, hist);
    hipLaunchKernelGGL((lic_eq_tile_hist<T, true>), dim3(tiles), dim3(kThreads), 0, stream, img,
                       ny, nx, sy, sx, nbins, st, tg, tcdf);
  } else {
    hipLaunchKernelGGL((lic_eq_hist<T, false>), dim3(gh), dim3(kThreads), 0, stream, img, n, nx,
                       sy, sx, nbins, st, hist);
    const int64_t counters = static_cast<int64_t>(tiles) * nbins;
    hipLaunchKernelGGL(lic_eq_zero, dim3(grid_for(counters, 8)), dim3(kThreads), 0, stream, tcdf,
                       counters);
    hipLaunchKernelGGL((lic_eq_tile_hist<T, false>), dim3(tiles), dim3(kThreads), 0, stream, img,
                       ny, nx, sy, sx, nbins, st, tg, tcdf);
  }
  hipLaunchKernelGGL(lic_eq_scan, dim3(1), dim3(kScanThreads), 0, stream, hist, nbins);
  hipLaunchKernelGGL(lic_eq_scan, dim3(tiles), dim3(kScanThreads), 0, stream, tcdf, nbins);
  hipLaunchKernelGGL(lic_eq_tile_map<T>, dim3(g), dim3(kThreads), 0, stream, img, ny, nx, sy, sx,
                     nbins, st, hist, tcdf, tg, out);
  return hipGetLastError();
}

template size_t equalize_tiles_scratch_bytes<float>(int, int32_t, int32_t, int32_t, int32_t,
                                                    int32_t);
template size_t equalize_tiles_scratch_bytes<double>(int, int32_t, int32_t, int32_t, int32_t,
                                                     int32_t);
template hipError_t launch_equalize_tiles<float>(int, const float*, int64_t, int64_t, int32_t,
                                                 int32_t, int32_t, int32_t, int32_t, float*,
                                                 void*, int32_t, hipStream_t);
template hipError_t launch_equalize_tiles<double>(int, const double*, int64_t, int64_t, int32_t,
                                                  int32_t, int32_t, int32_t, int32_t, double*,
                                                  void*, int32_t, hipStream_t);
template size_t equalize_scratch_bytes<float>(int32_t);
template size_t equalize_scratch_bytes<double>(int32_t);
template hipError_t launch_equalize<float>(const float*, int64_t, int64_t, int32_t, int32_t,
                                           int32_t, float*, void*, hipStream_t);
template hipError_t launch_equalize<double>(const double*, int64_t, int64_t, int32_t, int32_t,
                                            int32_t, double*, void*, hipStream_t);

}  // namespace rlic
