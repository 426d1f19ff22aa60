// Batched convolution (lic_batch.hip): host entry used by the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/rlic_hip.h"

namespace rlic {

// Stream-ordered scratch block owned by one batch call.
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

// Runs every field of the batch (already validated) in one persistent
// launch; synchronizes `stream` before returning.
template <typename T, typename FieldSpec>
int batch_run(const FieldSpec* specs, int64_t nfields, int profile, hipStream_t stream,
              BatchScratch& scratch, std::string& err);

}  // namespace rlic
