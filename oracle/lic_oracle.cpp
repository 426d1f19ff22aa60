// CPU restatement of rLIC's line integral convolution (the `rlic.convolve`
// numeric core, reference `src/lib.rs`).
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle: tests, the smoke check
// and the CPU-baseline leg of the benchmark load it as the checker.  The
// product path (rlic_amd) never links, loads or calls it.
//
// Each function cites the reference line range it restates.  Arithmetic is
// kept literal: explicit std::fma where the reference calls `mul_add`,
// unfused products everywhere else (build with -ffp-contract=off), IEEE
// division, no flush-to-zero.  Both Cargo feature flags of the reference
// (`fma`, `branchless`; Cargo.toml:27-30) are template parameters so that all
// four arithmetic profiles can be reproduced.
//
// Coordinates are kept as signed 64-bit integers; the reference's
// `usize::wrapping_sub(1)` sentinel for "stepped off the left/top edge" is -1
// here (src/lib.rs:221, 84).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

enum { BOUND_CLOSED = 0, BOUND_PERIODIC = 1 };
enum { MODE_VELOCITY = 0, MODE_POLARIZATION = 1 };

// num-traits 0.2.19 `Signed::signum` for floats (src/lib.rs:12, 175-177):
// -1 when the sign bit is set (including -0.0 and -inf), +1 otherwise,
// NaN for NaN.
template <typename T>
inline T signum_ref(T x) {
  if (std::isnan(x)) return x;
  return std::signbit(x) ? T(-1) : T(1);
}

// src/lib.rs:157-180  time_to_next_pixel
template <typename T, bool FMA, bool BRANCHLESS>
inline T time_to_next_pixel(T velocity, T frac) {
  if (!BRANCHLESS) {
    // src/lib.rs:158-166
    if (velocity > T(0)) return (T(1) - frac) / velocity;
    if (velocity < T(0)) return -(frac / velocity);
    return std::numeric_limits<T>::infinity();
  } else {
    // src/lib.rs:168-179
    const T factor = T(1) + signum_ref(velocity);
    const T half_minus = T(0.5) - frac;
    T remaining;
    if (FMA) {
      remaining = std::fma(factor, half_minus, frac);  // :177
    } else {
      const T prod = factor * half_minus;  // :175, two roundings
      remaining = prod + frac;
    }
    return std::fabs(remaining / velocity);  // :178
  }
}

// `a.mul_add(b, c)` with the fma feature, `c += a * b` without it
// (src/lib.rs:227/231, 355/359, 378/382).
template <typename T, bool FMA>
inline T mul_acc(T a, T b, T c) {
  if (FMA) return std::fma(a, b, c);
  const T prod = a * b;
  return c + prod;
}

struct Bounds {
  int x_left, x_right, y_left, y_right;
  int64_t nx, ny;
};

// src/lib.rs:83-95  BoundaryPair::apply_one_dir, with -1 as the left sentinel.
inline void apply_one_dir(int64_t& c, int left, int right, int64_t size) {
  if (c == -1) {
    c = (left == BOUND_CLOSED) ? 0 : size - 1;
  } else if (c == size) {
    c = (right == BOUND_CLOSED) ? size - 1 : 0;
  }
}

// src/lib.rs:208-233  update_state
template <typename T, bool FMA>
inline void update_state(T vel_par, T vel_orth, int64_t& coord_par, T& frac_par,
                         T& frac_orth, T t_par) {
  if (vel_par >= T(0)) {  // true for -0.0 (:217)
    coord_par += 1;
    frac_par = T(0);
  } else {
    coord_par -= 1;  // wrapping_sub(1); 0 becomes the -1 sentinel (:221)
    frac_par = T(1);
  }
  frac_orth = mul_acc<T, FMA>(t_par, vel_orth, frac_orth);  // :227 / :231
}

// src/lib.rs:235-273  advance
template <typename T, bool FMA, bool BRANCHLESS>
inline void advance(T u, T v, int64_t& i, int64_t& j, T& fx, T& fy,
                    const Bounds& b) {
  if (u == T(0) && v == T(0)) return;  // :242-244 (either sign of zero)
  const T tx = time_to_next_pixel<T, FMA, BRANCHLESS>(u, fx);
  const T ty = time_to_next_pixel<T, FMA, BRANCHLESS>(v, fy);
  if (tx < ty) {  // :249; ties and NaN take the y branch
    update_state<T, FMA>(u, v, j, fx, fy, tx);
  } else {
    update_state<T, FMA>(v, u, i, fy, fx, ty);
  }
  // :270-272  BoundarySet::apply -> x on j, then y on i (:112-115)
  apply_one_dir(j, b.x_left, b.x_right, b.nx);
  apply_one_dir(i, b.y_left, b.y_right, b.ny);
}

struct View2 {
  const void* base;
  int64_t sy, sx;  // element strides
};

template <typename T>
inline T at(const View2& a, int64_t i, int64_t j) {
  return static_cast<const T*>(a.base)[i * a.sy + j * a.sx];
}

// src/lib.rs:304-362  convole_single_pixel (one direction)
template <typename T, bool FMA, bool BRANCHLESS, bool FORWARD>
inline void convolve_one_direction(T& acc, int64_t i0, int64_t j0,
                                   const View2& u, const View2& v,
                                   int uv_mode, const T* kernel, int64_t klen,
                                   const T* input, const Bounds& b) {
  int64_t i = i0, j = j0;
  T fx = T(0.5), fy = T(0.5);       // :315-318
  T last_u = T(0), last_v = T(0);   // :320-323
  const int64_t kmid = klen / 2;    // :325
  // :326-329 forward k = kmid+1 .. klen-1, backward k = kmid-1 .. 0
  const int64_t k_begin = FORWARD ? kmid + 1 : kmid - 1;
  const int64_t k_end = FORWARD ? klen : -1;
  const int64_t k_step = FORWARD ? 1 : -1;
  for (int64_t k = k_begin; k != k_end; k += k_step) {
    T pu = at<T>(u, i, j);  // :332-335
    T pv = at<T>(v, i, j);
    if (std::isnan(pu) || std::isnan(pv)) break;  // :336-338
    if (uv_mode == MODE_POLARIZATION) {           // :339-347
      const T d1 = pu * last_u;  // unfused dot product (:341)
      const T d2 = pv * last_v;
      if (d1 + d2 < T(0)) {
        pu = -pu;
        pv = -pv;
      }
      last_u = pu;
      last_v = pv;
    }
    const T mu = FORWARD ? pu : -pu;  // :348-351
    const T mv = FORWARD ? pv : -pv;
    advance<T, FMA, BRANCHLESS>(mu, mv, i, j, fx, fy, b);  // :352
    acc = mul_acc<T, FMA>(kernel[k], input[i * b.nx + j], acc);  // :353-360
  }
}

// src/lib.rs:364-406  convolve (one pass over the image), rows [row0, row1)
template <typename T, bool FMA, bool BRANCHLESS>
void convolve_pass(const View2& u, const View2& v, int uv_mode,
                   const T* kernel, int64_t klen, const T* input, T* output,
                   const Bounds& b, int64_t row0, int64_t row1, int threads) {
  const int64_t kmid = klen / 2;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) if (threads > 1)
#endif
  for (int64_t i = row0; i < row1; ++i) {
    for (int64_t j = 0; j < b.nx; ++j) {
      T acc = T(0);  // output starts zeroed (:425, :438)
      acc = mul_acc<T, FMA>(kernel[kmid], input[i * b.nx + j], acc);  // :376-383
      convolve_one_direction<T, FMA, BRANCHLESS, true>(
          acc, i, j, u, v, uv_mode, kernel, klen, input, b);
      convolve_one_direction<T, FMA, BRANCHLESS, false>(
          acc, i, j, u, v, uv_mode, kernel, klen, input, b);
      output[i * b.nx + j] = acc;
    }
  }
  (void)threads;
}

// src/lib.rs:408-443  convolve_iteratively.  `work` holds ny*nx elements of
// scratch; `out` receives the result.  Only rows [row0, row1) of the final
// output are computed (used for bounded CPU-baseline samples; callers doing
// a full convolution pass 0 and ny).
template <typename T, bool FMA, bool BRANCHLESS>
int convolve_iter(const T* tex, int64_t tsy, int64_t tsx, const T* u,
                  int64_t usy, int64_t usx, const T* v, int64_t vsy,
                  int64_t vsx, int64_t ny, int64_t nx, const T* kernel,
                  int64_t klen, int uv_mode, const int* bounds,
                  int64_t iterations, T* out, T* work, int64_t row0,
                  int64_t row1, int threads) {
  if (ny < 0 || nx < 0 || klen < 0 || iterations < 0) return -1;
  const int64_t n = ny * nx;
  if (n == 0) return 0;
  if (iterations == 0) {
    for (int64_t i = 0; i < ny; ++i)
      for (int64_t j = 0; j < nx; ++j) out[i * nx + j] = tex[i * tsy + j * tsx];
    return 0;
  }
  if (klen == 0) return -2;  // the reference aborts on kernel[[kmid]] here
  Bounds b{bounds[0], bounds[1], bounds[2], bounds[3], nx, ny};
  View2 uu{u, usy, usx}, vv{v, vsy, vsx};
  // :423-424 C-order copy of the texture into the iteration input
  T* input = work;
  for (int64_t i = 0; i < ny; ++i)
    for (int64_t j = 0; j < nx; ++j) input[i * nx + j] = tex[i * tsy + j * tsx];
  for (int64_t it = 0; it < iterations; ++it) {  // :432-440
    const bool last = (it + 1 == iterations);
    // intermediate passes must cover every row; the last may be banded
    const int64_t r0 = last ? row0 : 0, r1 = last ? row1 : ny;
    convolve_pass<T, FMA, BRANCHLESS>(uu, vv, uv_mode, kernel, klen, input, out,
                                     b, r0, r1, threads);
    if (!last) std::memcpy(input, out, sizeof(T) * n);  // input.assign(&output)
  }
  return 0;
}

template <typename T>
int dispatch(int profile, const T* tex, int64_t tsy, int64_t tsx, const T* u,
             int64_t usy, int64_t usx, const T* v, int64_t vsy, int64_t vsx,
             int64_t ny, int64_t nx, const T* kernel, int64_t klen,
             int uv_mode, const int* bounds, int64_t iterations, T* out,
             T* work, int64_t row0, int64_t row1, int threads) {
  switch (profile & 3) {
    case 0:
      return convolve_iter<T, false, false>(tex, tsy, tsx, u, usy, usx, v, vsy,
                                            vsx, ny, nx, kernel, klen, uv_mode,
                                            bounds, iterations, out, work,
                                            row0, row1, threads);
    case 1:
      return convolve_iter<T, true, false>(tex, tsy, tsx, u, usy, usx, v, vsy,
                                           vsx, ny, nx, kernel, klen, uv_mode,
                                           bounds, iterations, out, work, row0,
                                           row1, threads);
    case 2:
      return convolve_iter<T, false, true>(tex, tsy, tsx, u, usy, usx, v, vsy,
                                           vsx, ny, nx, kernel, klen, uv_mode,
                                           bounds, iterations, out, work, row0,
                                           row1, threads);
    default:
      return convolve_iter<T, true, true>(tex, tsy, tsx, u, usy, usx, v, vsy,
                                          vsx, ny, nx, kernel, klen, uv_mode,
                                          bounds, iterations, out, work, row0,
                                          row1, threads);
  }
}

template <typename T>
T ttnp_dispatch(int profile, T vel, T frac) {
  switch (profile & 3) {
    case 0: return time_to_next_pixel<T, false, false>(vel, frac);
    case 1: return time_to_next_pixel<T, true, false>(vel, frac);
    case 2: return time_to_next_pixel<T, false, true>(vel, frac);
    default: return time_to_next_pixel<T, true, true>(vel, frac);
  }
}

template <typename T>
void advance_dispatch(int profile, T u, T v, int64_t* i, int64_t* j, T* fx,
                      T* fy, int64_t ny, int64_t nx, const int* bounds) {
  Bounds b{bounds[0], bounds[1], bounds[2], bounds[3], nx, ny};
  switch (profile & 3) {
    case 0: advance<T, false, false>(u, v, *i, *j, *fx, *fy, b); break;
    case 1: advance<T, true, false>(u, v, *i, *j, *fx, *fy, b); break;
    case 2: advance<T, false, true>(u, v, *i, *j, *fx, *fy, b); break;
    default: advance<T, true, true>(u, v, *i, *j, *fx, *fy, b); break;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// C entry points (loaded with ctypes by the tests and the benchmark).
//
// profile: bit 0 = the reference's `fma` feature, bit 1 = `branchless`.
//   3 = Cargo default (fma + branchless), 1 = fma only (x86_64 wheels).
// bounds: {x_left, x_right, y_left, y_right}, 0 = closed, 1 = periodic.
// Strides are in elements.  `work` must hold ny*nx elements.
// rows [row0, row1) of the final pass are computed (0, ny for everything).
// Returns 0, -1 on a bad size, -2 for an empty kernel with a non-empty image.
// ---------------------------------------------------------------------------
extern "C" {

int lic_oracle_convolve_f64(const double* tex, int64_t tsy, int64_t tsx,
                            const double* u, int64_t usy, int64_t usx,
                            const double* v, int64_t vsy, int64_t vsx,
                            int64_t ny, int64_t nx, const double* kernel,
                            int64_t klen, int uv_mode, const int* bounds,
                            int64_t iterations, int profile, double* out,
                            double* work, int64_t row0, int64_t row1,
                            int threads) {
  return dispatch<double>(profile, tex, tsy, tsx, u, usy, usx, v, vsy, vsx, ny,
                          nx, kernel, klen, uv_mode, bounds, iterations, out,
                          work, row0, row1, threads);
}

int lic_oracle_convolve_f32(const float* tex, int64_t tsy, int64_t tsx,
                            const float* u, int64_t usy, int64_t usx,
                            const float* v, int64_t vsy, int64_t vsx,
                            int64_t ny, int64_t nx, const float* kernel,
                            int64_t klen, int uv_mode, const int* bounds,
                            int64_t iterations, int profile, float* out,
                            float* work, int64_t row0, int64_t row1,
                            int threads) {
  return dispatch<float>(profile, tex, tsy, tsx, u, usy, usx, v, vsy, vsx, ny,
                         nx, kernel, klen, uv_mode, bounds, iterations, out,
                         work, row0, row1, threads);
}

double lic_oracle_time_to_next_pixel_f64(double vel, double frac, int profile) {
  return ttnp_dispatch<double>(profile, vel, frac);
}

float lic_oracle_time_to_next_pixel_f32(float vel, float frac, int profile) {
  return ttnp_dispatch<float>(profile, vel, frac);
}

void lic_oracle_advance_f64(double u, double v, int64_t* i, int64_t* j,
                            double* fx, double* fy, int64_t ny, int64_t nx,
                            const int* bounds, int profile) {
  advance_dispatch<double>(profile, u, v, i, j, fx, fy, ny, nx, bounds);
}

void lic_oracle_advance_f32(float u, float v, int64_t* i, int64_t* j,
                            float* fx, float* fy, int64_t ny, int64_t nx,
                            const int* bounds, int profile) {
  advance_dispatch<float>(profile, u, v, i, j, fx, fy, ny, nx, bounds);
}

int lic_oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
