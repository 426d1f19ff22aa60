// Work-queue launches over many fields (lic_batch.hip), shared by the
// batched call and the serving step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/rlic_hip.h"
#include "lic_device.h"

namespace rlic {

// Stream-ordered device scratch owned by one call.
struct BatchScratch {
  void* ptr = nullptr;
  hipStream_t stream = nullptr;
  BatchScratch() = default;
  BatchScratch(const BatchScratch&) = delete;
  BatchScratch& operator=(const BatchScratch&) = delete;
  ~BatchScratch() {
    if (ptr) (void)hipFreeAsync(ptr, stream);
  }
  hipError_t reserve(size_t bytes, hipStream_t s) {
    stream = s;
    return hipMallocAsync(&ptr, bytes ? bytes : 1, s);
  }
};

// One field of a queue launch, as the device reads it.
template <typename T>
struct BatchField {
  const T* tex;  // texture, u, v with their strides: read by the pack step only
  int64_t tsy, tsx;
  const T* u;
  int64_t usy, usx;
  const T* v;
  int64_t vsy, vsx;
  const T* kernel;      // set by launch_queue: first weight of the padded copy
  int64_t kernel_off;   // set by launch_queue: that weight as an offset into the pool
  T* out;               // pack step: receives the texture when iterations == 0
  Cell<T>* cells;       // packed (u, v, texture)
  uint4* codes;         // move record, or nullptr
  FieldGeom g;
  int32_t klen, uv_mode;
  int32_t iterations;   // pack step only (0: copy the texture to out)
  int32_t regions_x, regions;  // set by launch_queue
  int64_t pack_block0;         // set by launch_queue
};

// A run of queue items: every region of one pass of one field.
struct Segment {
  int64_t item0;    // set by launch_queue
  int32_t field;
  int32_t pass;     // 0 traces from the cells, > 0 replays the move record
  int32_t prev;     // segment whose items must all be done first (-1: none)
  int32_t record;   // pass 0: write the move record
  const void* src;  // replay input: the previous pass's output
  void* dst;        // this pass's output
};

// Pads every field's kernel weights (kernel_src[k], device memory) into a
// weight pool, packs every field's (texture, u, v) into its cells when
// `pack`, then runs the items of every segment in one persistent launch.
// With `deps`, the items of a segment wait for segment `prev`, which must
// come earlier in `segs`; without, no segment may have a `prev`.  Fills the
// fields' and segments' launch_queue members and synchronizes `stream`.
template <typename T>
int launch_queue(std::vector<BatchField<T>>& fields, const std::vector<const T*>& kernel_src,
                 std::vector<Segment>& segs, bool deps, bool pack, int profile,
                 hipStream_t stream, std::string& err);

// The element checks of rlic.convolve (src/rlic/_lib.py:174, :194) for
// one field or every field of a batch at once, on the device.
template <typename T>
struct CheckField {
  const T* texture;
  int64_t tex_sy, tex_sx, ny, nx;
  const T* kernel;  // contiguous
  int64_t kernel_len;
};
constexpr int32_t kCheckInline = 32;
template <typename T>
struct CheckPack {
  CheckField<T> f[kCheckInline];
};

// flags[2k] (texture element below 0) and flags[2k + 1] (non-finite
// weight) of field k become 1; the caller zeroes them first (they may be
// host-mapped memory).  Fields from `table` (device memory) or, for up to
// kCheckInline fields, from `pack` (copied into the kernel arguments).
// `max_pixels`: the largest texture's element count (sizes the grid).
template <typename T>
hipError_t launch_check_values(const CheckField<T>* table, const CheckPack<T>* pack,
                               int32_t nfields, int64_t max_pixels, uint32_t* flags,
                               hipStream_t stream);

// Runs every field of the batch (already validated) in one persistent
// launch; synchronizes `stream` before returning.
template <typename T, typename FieldSpec>
int batch_run(const FieldSpec* specs, int64_t nfields, int profile, hipStream_t stream,
              BatchScratch& scratch, std::string& err);

}  // namespace rlic
