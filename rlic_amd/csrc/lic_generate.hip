// On-device synthetic inputs: uniform noise textures and Fourier-mode vector
// fields written straight into device memory, so that large benchmark and
// serving inputs need no host generation or host-to-device copy.  Neither
// has a counterpart in the reference (its inputs are NumPy arrays).
//
//   lic_fill_noise  out[p] = uniform in [0, 1) from a counter-based hash of
//                   (seed, p): splitmix64 (Steele, Lea and Flood, "Fast
//                   splittable pseudorandom number generators", OOPSLA 2014)
//                   of seed + (p + 1) * 0x9E3779B97F4A7C15, top 53 bits
//                   (double) or 24 bits (float) as a binary fraction
//   lic_fill_modes  out[i, j] = sum over modes m of
//                     a_m (cos(ay) cos(ax) - sin(ay) sin(ax)),
//                   ax = 2 pi kx_m (j / nx) + phi_m, ay = 2 pi ky_m (i / ny),
//                   evaluated in double in that order, then rounded to T
#include "lic_launch.h"

namespace rlic {

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <typename T>
__device__ __forceinline__ T unit_fraction(uint64_t z);

template <>
__device__ __forceinline__ double unit_fraction<double>(uint64_t z) {
  return static_cast<double>(z >> 11) * 0x1.0p-53;
}

template <>
__device__ __forceinline__ float unit_fraction<float>(uint64_t z) {
  return static_cast<float>(static_cast<uint32_t>(z >> 40)) * 0x1.0p-24f;
}

template <typename T>
__global__ __launch_bounds__(256) void lic_fill_noise(T* __restrict__ out, int64_t n,
                                                      uint64_t seed) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * 256;
  for (int64_t p = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; p < n; p += stride)
    out[p] = unit_fraction<T>(splitmix64(seed + static_cast<uint64_t>(p + 1) * 0x9E3779B97F4A7C15ull));
}

template <typename T>
__global__ __launch_bounds__(256) void lic_fill_modes(T* __restrict__ out, int32_t ny, int32_t nx,
                                                      ModeSet modes) {
  const int32_t j = blockIdx.x * 64 + (threadIdx.x & 63);
  if (j >= nx) return;
  const double two_pi = 6.283185307179586;  // 2 * pi rounded to double
  const double jj = static_cast<double>(j) / nx;
  for (int32_t i = blockIdx.y * 4 + (threadIdx.x >> 6); i < ny; i += gridDim.y * 4) {
    const double ii = static_cast<double>(i) / ny;
    double acc = 0.0;
    for (int32_t m = 0; m < modes.n; ++m) {
      const double ax = __dadd_rn(__dmul_rn(__dmul_rn(two_pi, modes.kx[m]), jj), modes.phase[m]);
      const double ay = __dmul_rn(__dmul_rn(two_pi, modes.ky[m]), ii);
      const double c = __dsub_rn(__dmul_rn(cos(ay), cos(ax)), __dmul_rn(sin(ay), sin(ax)));
      acc = __dadd_rn(acc, __dmul_rn(modes.amp[m], c));
    }
    out[static_cast<int64_t>(i) * nx + j] = static_cast<T>(acc);
  }
}

template <typename T>
hipError_t launch_fill_noise(T* out, int64_t n, uint64_t seed, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const int64_t blocks = (n + 255) / 256;
  const unsigned grid = static_cast<unsigned>(blocks < 65536 ? blocks : 65536);
  hipLaunchKernelGGL(lic_fill_noise<T>, dim3(grid), dim3(256), 0, stream, out, n, seed);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_fill_modes(T* out, int32_t ny, int32_t nx, const ModeSet& modes,
                             hipStream_t stream) {
  if (ny == 0 || nx == 0) return hipSuccess;
  const int32_t rows = (ny + 3) / 4;
  const dim3 grid((nx + 63) / 64, rows < 65535 ? rows : 65535);
  hipLaunchKernelGGL(lic_fill_modes<T>, grid, dim3(256), 0, stream, out, ny, nx, modes);
  return hipGetLastError();
}

template hipError_t launch_fill_noise<float>(float*, int64_t, uint64_t, hipStream_t);
template hipError_t launch_fill_noise<double>(double*, int64_t, uint64_t, hipStream_t);
template hipError_t launch_fill_modes<float>(float*, int32_t, int32_t, const ModeSet&, hipStream_t);
template hipError_t launch_fill_modes<double>(double*, int32_t, int32_t, const ModeSet&,
                                              hipStream_t);

}  // namespace rlic
