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

// Inclusive scan of the counts, in place (one block of kScanThreads).
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void lic_eq_scan(uint32_t* hist, int32_t nbins) {
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

template size_t equalize_scratch_bytes<float>(int32_t);
template size_t equalize_scratch_bytes<double>(int32_t);
template hipError_t launch_equalize<float>(const float*, int64_t, int64_t, int32_t, int32_t,
                                           int32_t, float*, void*, hipStream_t);
template hipError_t launch_equalize<double>(const double*, int64_t, int64_t, int32_t, int32_t,
                                            int32_t, double*, void*, hipStream_t);

}  // namespace rlic
