/*
 * rlic_hip.h — C ABI of the MI355X (gfx950) line integral convolution library
 * librlic_hip.so.
 *
 * The entry points replace the reference's PyO3 functions
 *   rlic._core.convolve_f32 / convolve_f64      (src/lib.rs:450-482,
 *                                                src/rlic/_core.pyi:8-29)
 * which the Python API calls at exactly one site (src/rlic/_lib.py:235).
 * Arguments map one to one:
 *   texture                    -> texture + element strides (sy, sx)
 *   uv = (u, v, uv_mode)       -> u, v + strides, uv_mode (RLIC_UV_*)
 *   kernel                     -> kernel, kernel_len
 *   boundaries ((xl,xr),(yl,yr)) -> x_left, x_right, y_left, y_right (RLIC_BOUND_*)
 *   iterations                 -> iterations
 *   return value (new ndarray) -> caller-allocated C-order `out` (ny*nx)
 *
 * Conventions
 *   - Plain pointers and sizes only.  Strides are in ELEMENTS and may be
 *     negative or zero (broadcast views); element (i, j) of an array is at
 *     base[i*sy + j*sx].  `out` is always C-contiguous and must not overlap
 *     any input.
 *   - `location` says where ALL array pointers live: RLIC_HOST (pageable or
 *     pinned host memory; the call uploads, computes, downloads and returns
 *     when `out` is filled) or RLIC_DEVICE (device memory of the current HIP
 *     device; work is enqueued on `stream` and the call returns without
 *     waiting, like any stream-ordered HIP operation).
 *   - `stream` is a hipStream_t (NULL = the legacy default stream).
 *   - Return 0 on success or a negative RLIC_ERR_* code; nothing aborts.
 *     rlic_last_error() gives a message for the calling thread's last failure.
 *   - Thread-safe.  The library keeps no pointer past return.
 *
 * profile (bit flags) selects the reference's compile-time arithmetic
 * profile (Cargo features, Cargo.toml:27-30):
 *   RLIC_PROFILE_FMA        `fma`        mul_add at src/lib.rs:177,231,359,382
 *   RLIC_PROFILE_BRANCHLESS `branchless` time_to_next_pixel at src/lib.rs:168-179
 *   RLIC_PROFILE_DEFAULT = both, the Cargo default (Cargo.toml:28).
 */
#ifndef RLIC_HIP_H
#define RLIC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  RLIC_UV_VELOCITY = 0,      /* src/lib.rs:16  */
  RLIC_UV_POLARIZATION = 1   /* src/lib.rs:17  */
};

enum {
  RLIC_BOUND_CLOSED = 0,     /* src/lib.rs:57 */
  RLIC_BOUND_PERIODIC = 1    /* src/lib.rs:58 */
};

enum {
  RLIC_HOST = 0,
  RLIC_DEVICE = 1
};

enum {
  RLIC_PROFILE_FMA = 1,
  RLIC_PROFILE_BRANCHLESS = 2,
  RLIC_PROFILE_DEFAULT = 3
};

enum {
  RLIC_OK = 0,
  RLIC_ERR_INVALID_ARGUMENT = -1,
  RLIC_ERR_OUT_OF_MEMORY = -2,
  RLIC_ERR_HIP = -3,
  RLIC_ERR_EMPTY_KERNEL = -4  /* the reference aborts the process here
                                 (kernel[[kmid]] out of bounds, src/lib.rs:382) */
};

/* Replaces rlic._core.convolve_f32 (src/lib.rs:450-465). */
int rlic_convolve_f32(const float* texture, int64_t tex_sy, int64_t tex_sx,
                      const float* u, int64_t u_sy, int64_t u_sx,
                      const float* v, int64_t v_sy, int64_t v_sx,
                      int64_t ny, int64_t nx,
                      const float* kernel, int64_t kernel_len,
                      int uv_mode,
                      int x_left, int x_right, int y_left, int y_right,
                      int64_t iterations,
                      float* out,
                      int location, int profile, void* stream);

/* Replaces rlic._core.convolve_f64 (src/lib.rs:467-482). */
int rlic_convolve_f64(const double* texture, int64_t tex_sy, int64_t tex_sx,
                      const double* u, int64_t u_sy, int64_t u_sx,
                      const double* v, int64_t v_sy, int64_t v_sx,
                      int64_t ny, int64_t nx,
                      const double* kernel, int64_t kernel_len,
                      int uv_mode,
                      int x_left, int x_right, int y_left, int y_right,
                      int64_t iterations,
                      double* out,
                      int location, int profile, void* stream);

/* Allocation-free variant of rlic_convolve_f32 / _f64 for device arrays.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, checks and results as
 * rlic_convolve_* with location = RLIC_DEVICE (enqueued on `stream`;
 * returns without waiting).  The call's scratch (padded weights, guarded
 * cells, ping-pong buffer, move record) lives in one fixed device array of
 * RLIC_NOALLOC_SCRATCH_BYTES that each GPU holds from its first such call
 * until the process ends.  A call whose weights, cells and ping-pong buffer
 * do not fit fails with RLIC_ERR_INVALID_ARGUMENT before any work; one whose
 * move record does not also fit traces every pass instead of replaying
 * (same bits, more work).  No hipMalloc, hipFree, events or graphs.
 * Every call on a GPU shares that GPU's array and leaves its contents
 * there: calls on one stream are ordered, but calls that overlap on one GPU
 * (other streams or other host threads) race, and the caller must
 * serialize them. */
#define RLIC_NOALLOC_SCRATCH_BYTES (64ll << 20)
int rlic_convolve_noalloc_f32(const float* texture, int64_t tex_sy, int64_t tex_sx,
                              const float* u, int64_t u_sy, int64_t u_sx,
                              const float* v, int64_t v_sy, int64_t v_sx,
                              int64_t ny, int64_t nx,
                              const float* kernel, int64_t kernel_len,
                              int uv_mode,
                              int x_left, int x_right, int y_left, int y_right,
                              int64_t iterations,
                              float* out,
                              int profile, void* stream);
int rlic_convolve_noalloc_f64(const double* texture, int64_t tex_sy, int64_t tex_sx,
                              const double* u, int64_t u_sy, int64_t u_sx,
                              const double* v, int64_t v_sy, int64_t v_sx,
                              int64_t ny, int64_t nx,
                              const double* kernel, int64_t kernel_len,
                              int uv_mode,
                              int x_left, int x_right, int y_left, int y_right,
                              int64_t iterations,
                              double* out,
                              int profile, void* stream);

/* Global histogram equalization of an image: the step usually applied to a
 * LIC image, run where that image lies (with RLIC_DEVICE, on the output of
 * a device rlic_convolve_* call, which never leaves HBM).  Replaces the
 * reference's declared but unimplemented rlic._core.equalize_histogram_f32
 * / _f64(image, nbins) (src/rlic/_core.pyi:30-37; README.md:19-23 points
 * to the sister project `ahe`).  out(x) = cdf(bin(x)) / N over the finite
 * pixels: vmin, vmax the finite range, bin(x) = min(floor((x - vmin) *
 * (nbins / (vmax - vmin))), nbins - 1) in double (bin 0 when vmax == vmin),
 * N the finite pixel count; NaN and +-inf pixels are copied unchanged.
 * `image` has element strides (sy, sx); `out` is C-order ny x nx and may be
 * `image` itself when that is C-contiguous.  nbins in [1, 2^24].  Same
 * `location` / `stream` conventions as rlic_convolve_*. */
int rlic_equalize_histogram_f32(const float* image, int64_t sy, int64_t sx, int64_t ny,
                                int64_t nx, int64_t nbins, float* out, int location,
                                void* stream);
int rlic_equalize_histogram_f64(const double* image, int64_t sy, int64_t sx, int64_t ny,
                                int64_t nx, int64_t nbins, double* out, int location,
                                void* stream);

/* Adaptive histogram equalization, replacing the reference's declared but
 * unimplemented rlic._core.equalize_histogram_sliding_tile_f32/_f64 and
 * equalize_histogram_tile_interpolation_f32/_f64(image, nbins, tile_shape)
 * (src/rlic/_core.pyi:38-57).  Bins as in rlic_equalize_histogram_*
 * (over the whole image's finite range); NaN and +-inf pixels copied.
 * Sliding tile: out(p) = (finite pixels of p's window with a bin <= p's) /
 * (finite pixels of the window), the window rows
 * [i - (tile_h - 1) / 2, i - (tile_h - 1) / 2 + tile_h) and columns alike,
 * clipped to the image (cost: tile_h * tile_w per pixel).  Tile
 * interpolation: tiles of tile_h x tile_w from (0, 0), tile t mapping bin b
 * to cdf_t(b) / N_t (the global mapping for a tile without finite pixels),
 * bilinear interpolation between the mappings of the tiles whose centres
 * surround the pixel (clamped at the edges), in double.  tile_h, tile_w
 * >= 1; for tile interpolation tiles x nbins <= 2^30.  Other arguments as
 * rlic_equalize_histogram_*. */
int rlic_equalize_histogram_sliding_tile_f32(const float* image, int64_t sy, int64_t sx,
                                             int64_t ny, int64_t nx, int64_t nbins,
                                             int64_t tile_h, int64_t tile_w, float* out,
                                             int location, void* stream);
int rlic_equalize_histogram_sliding_tile_f64(const double* image, int64_t sy, int64_t sx,
                                             int64_t ny, int64_t nx, int64_t nbins,
                                             int64_t tile_h, int64_t tile_w, double* out,
                                             int location, void* stream);
int rlic_equalize_histogram_tile_interpolation_f32(const float* image, int64_t sy, int64_t sx,
                                                   int64_t ny, int64_t nx, int64_t nbins,
                                                   int64_t tile_h, int64_t tile_w, float* out,
                                                   int location, void* stream);
int rlic_equalize_histogram_tile_interpolation_f64(const double* image, int64_t sy, int64_t sx,
                                                   int64_t ny, int64_t nx, int64_t nbins,
                                                   int64_t tile_h, int64_t tile_w, double* out,
                                                   int location, void* stream);

/* One field of a batched call: the arguments of one rlic_convolve_* call
 * with DEVICE pointers (see rlic_convolve_batch_*). */
typedef struct {
  const float* texture; int64_t tex_sy, tex_sx;
  const float* u; int64_t u_sy, u_sx;
  const float* v; int64_t v_sy, v_sx;
  int64_t ny, nx;
  const float* kernel; int64_t kernel_len;
  int uv_mode, x_left, x_right, y_left, y_right;
  int64_t iterations;
  float* out;
} rlic_field_f32;

typedef struct {
  const double* texture; int64_t tex_sy, tex_sx;
  const double* u; int64_t u_sy, u_sx;
  const double* v; int64_t v_sy, v_sx;
  int64_t ny, nx;
  const double* kernel; int64_t kernel_len;
  int uv_mode, x_left, x_right, y_left, y_right;
  int64_t iterations;
  double* out;
} rlic_field_f64;

/* Batched convolution: n_fields independent rlic_convolve_* calls (device
 * memory only).  Fields of at least 2^18 pixels (RLIC_BATCH_SINGLE_MIN)
 * run as single-field launches; the rest run together in one persistent
 * launch fed by a device-side work queue.  Results are the same, bit for
 * bit, as the single calls.  Returns after the work has completed on
 * `stream`.  No counterpart in the
 * reference (which has one call per field, src/lib.rs:450-482); each field
 * follows the semantics of that call. */
int rlic_convolve_batch_f32(const rlic_field_f32* fields, int64_t n_fields, int profile,
                            void* stream);
int rlic_convolve_batch_f64(const rlic_field_f64* fields, int64_t n_fields, int profile,
                            void* stream);

/* The element checks of rlic.convolve: a texture element below 0
 * (src/rlic/_lib.py:174) and a non-finite kernel weight (:194), which the
 * reference runs on the host with NumPy (`np.any(texture < 0)`,
 * `np.isfinite(kernel)`).  Here one device launch reads the arrays in
 * place (device memory, strided textures; a NaN texture element passes, as
 * in the reference) and writes its flags to host-mapped memory; the call
 * returns after that launch has completed on `stream`.
 * rlic_check_values_*: one field; *what = bit 0 set when the texture has
 * an element below 0, bit 1 when the kernel has a non-finite weight.
 * rlic_batch_check_values_*: every field of a batch (only the texture,
 * stride, size and kernel members are read); *first_bad = the index of the
 * first failing field or -1, *what (may be null) = that field's bits. */
int rlic_check_values_f32(const float* texture, int64_t ny, int64_t nx, int64_t tex_sy,
                          int64_t tex_sx, const float* kernel, int64_t kernel_len, int* what,
                          void* stream);
int rlic_check_values_f64(const double* texture, int64_t ny, int64_t nx, int64_t tex_sy,
                          int64_t tex_sx, const double* kernel, int64_t kernel_len, int* what,
                          void* stream);
int rlic_batch_check_values_f32(const rlic_field_f32* fields, int64_t n_fields,
                                int64_t* first_bad, int* what, void* stream);
int rlic_batch_check_values_f64(const rlic_field_f64* fields, int64_t n_fields,
                                int64_t* first_bad, int* what, void* stream);

/* Serving: fields resident in device memory, advanced one pass per step.
 * No counterpart in the reference (one blocking call per field,
 * src/lib.rs:450-482); every field's passes follow the semantics of that
 * call, and its output after n passes equals rlic_convolve_* with
 * iterations = n, bit for bit, whatever else shares the steps.
 *
 * rlic_server_create: a server on the current device that keeps at most
 *   budget_bytes of field state resident and runs its work on `stream`.
 *   NULL on failure (see rlic_last_error).
 * rlic_server_admit_*: copies the kernel weights and packs (u, v, texture)
 *   (device memory; `out` is ignored), registers the
 *   field for `iterations` (>= 1) passes and returns its handle.  keep != 0
 *   keeps the move record so that rlic_server_extend can add passes later.
 *   RLIC_ERR_OUT_OF_MEMORY when the field does not fit the free budget
 *   (nothing is admitted then; retry after releasing fields).
 * rlic_server_field_bytes_*: the budget bytes such a field occupies.
 * rlic_server_step: runs one pass of every field with passes left and
 *   writes the handles whose requested passes are now all done into
 *   finished[0 .. *n_finished) (capacity: at least the number of active
 *   fields).  Returns after the work has completed.
 * rlic_server_extend: requests `more` passes of a kept field whose passes
 *   are all done (a follow-up: it continues from the field's last output).
 * rlic_server_read_*: copies the field's latest output (ny x nx, C order)
 *   to `out` (device or host memory); synchronous.
 * rlic_server_release: frees the field's state and budget.
 * rlic_server_usage: budget bytes in use, resident and active fields.
 * Calls on one server are serialized; distinct servers are independent. */
typedef struct rlic_server rlic_server;
rlic_server* rlic_server_create(int64_t budget_bytes, int profile, void* stream);
void rlic_server_destroy(rlic_server* server);
int rlic_server_admit_f32(rlic_server* server, const rlic_field_f32* field, int keep,
                          int64_t* handle);
int rlic_server_admit_f64(rlic_server* server, const rlic_field_f64* field, int keep,
                          int64_t* handle);
int64_t rlic_server_field_bytes_f32(int64_t ny, int64_t nx, int64_t kernel_len, int keep,
                                    int64_t iterations);
int64_t rlic_server_field_bytes_f64(int64_t ny, int64_t nx, int64_t kernel_len, int keep,
                                    int64_t iterations);
int rlic_server_step(rlic_server* server, int64_t* finished, int64_t capacity,
                     int64_t* n_finished);
int rlic_server_extend(rlic_server* server, int64_t handle, int64_t more);
int rlic_server_read_f32(rlic_server* server, int64_t handle, float* out);
int rlic_server_read_f64(rlic_server* server, int64_t handle, double* out);
int rlic_server_release(rlic_server* server, int64_t handle);
int rlic_server_usage(rlic_server* server, int64_t* used_bytes, int64_t* resident,
                      int64_t* active);

/* Message for the calling thread's most recent failed call ("" if none).
 * No counterpart in the reference, whose errors abort (Cargo.toml:23). */
const char* rlic_last_error(void);

/* Kernel timing.  No counterpart in the reference (which has no timers).
 * rlic_profiler_enable(1): every later kernel launch of the library is
 * bracketed by two hipEvents recorded on the stream it runs on.
 * rlic_profiler_read: resolves the recorded events (waits for them),
 * writes a JSON object {"<kernel>": {"count": n, "total_ms": t}, ...} into
 * buf (NUL-terminated) and clears the record when len is large enough,
 * keeps the record otherwise (e.g. buf == NULL to size the buffer), and
 * returns the number of bytes needed including the NUL. */
int rlic_profiler_enable(int on);
int64_t rlic_profiler_read(char* buf, int64_t len);

/* Divergence measurement.  No counterpart in the reference (a serial loop,
 * src/lib.rs:373-374).  With the environment variable RLIC_TRACE_STATS=1,
 * every tracing pass of rlic_convolve_* also records its moves and reads
 * them back to count, over the lanes of its waves, the steps each lane's
 * streamline was alive and the steps its wave executed (a wave runs until
 * all its lanes' lines have ended).  rlic_trace_stats_read returns the two
 * sums accumulated since the last read and resets them; their ratio is the
 * lane efficiency lost to uneven streamline termination. */
int rlic_trace_stats_read(int64_t* live_lane_steps, int64_t* executed_lane_steps);

/* Replay autotuner.  No counterpart in the reference (one CPU loop).  The
 * first rlic_convolve_* call with at least 3 passes on at least 2^18 pixels
 * for a new (dtype, shape, kernel length, profile) times its second pass
 * under each candidate plan — LDS-staged replay with every tile whose box
 * (tile + kernel-half-length halo) fits the LDS, and L2-gather replay with
 * two tile strip heights — and later calls with that key use the fastest.
 * Every plan gives the same bits.  RLIC_AUTOTUNE=0 disables tuning;
 * RLIC_AUTOTUNE_CACHE=<file> loads the table from and appends it to a text
 * file.  rlic_autotune_read writes the table as a JSON list (same buffer
 * convention as rlic_profiler_read); rlic_autotune_load adds the plan lines
 * of such a file (returns how many, or a negative status); a loaded plan
 * that cannot run on this device is ignored; rlic_autotune_clear empties
 * the table. */
int64_t rlic_autotune_read(char* buf, int64_t len);
int64_t rlic_autotune_load(const char* path);
int rlic_autotune_clear(void);

/* Check of the float division the streamline kernels use for step times
 * (src/lib.rs:157-180 divide in IEEE arithmetic): q[i] = a[i] / b[i] for n
 * device-memory elements, computed on the device; synchronous.  No
 * counterpart in the reference; for tests. */
int rlic_div_check_f32(const float* a, const float* b, float* q, int64_t n, void* stream);

/* Scratch memory.  A call's device scratch (packed cells, move record,
 * ping-pong buffer, queue descriptors) is one block; blocks up to 256 MiB
 * are cached per (device, stream) and reused by later calls on that stream.
 * rlic_scratch_release synchronizes the devices concerned and frees the
 * cached blocks.  No counterpart in the reference. */
int rlic_scratch_release(void);

/* On-device synthetic inputs, so that large benchmark and serving inputs
 * need no host generation or host-to-device copy.  No counterpart in the
 * reference (its inputs are NumPy arrays).  `out` is ny x nx contiguous
 * elements of device memory; the calls only enqueue work on `stream`.
 *
 * rlic_fill_noise_*: out[p] = uniform in [0, 1) from a counter-based hash
 * (splitmix64) of `seed` and the element index p, so a seed gives the same
 * array on any device and at any launch shape.
 *
 * rlic_fill_modes_*: out[i][j] = sum over m < n_modes (at most 32) of
 * amp[m] * cos(2 pi (kx[m] j / nx + ky[m] i / ny) + phase[m]), evaluated in
 * double; amp, kx, ky and phase are host arrays of n_modes elements. */
int rlic_fill_noise_f32(float* out, int64_t ny, int64_t nx, uint64_t seed, void* stream);
int rlic_fill_noise_f64(double* out, int64_t ny, int64_t nx, uint64_t seed, void* stream);
int rlic_fill_modes_f32(float* out, int64_t ny, int64_t nx, int32_t n_modes, const double* amp,
                        const int32_t* kx, const int32_t* ky, const double* phase, void* stream);
int rlic_fill_modes_f64(double* out, int64_t ny, int64_t nx, int32_t n_modes, const double* amp,
                        const int32_t* kx, const int32_t* ky, const double* phase, void* stream);

/* Library version string.  No counterpart in the reference. */
const char* rlic_version(void);

#ifdef __cplusplus
}
#endif

#endif /* RLIC_HIP_H */
