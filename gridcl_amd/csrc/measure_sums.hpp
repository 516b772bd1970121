// Device helper of the measure's fixed-order sums (tetraloma-geometry/src/measure.rs:44-102):
// each pass adds blocks of 64 values in index order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tg {

constexpr uint32_t kSumBlock = 64;  // values per output of one pass (measure.rs:44)

// x[0] + x[1] + ... + x[count - 1] in order (0 < count <= 64), the values loaded kChunk at a
// time into registers before they are added, so the loads overlap.
template <typename T, uint32_t kChunk = sizeof(T) == 4 ? 64 : 32>
__device__ __forceinline__ T ordered_sum(const T* __restrict__ x, uint32_t count) {
  T acc = T(0);
#pragma unroll 1
  for (uint32_t c0 = 0; c0 < kSumBlock; c0 += kChunk) {
    T v[kChunk];
#pragma unroll
    for (uint32_t j = 0; j < kChunk; ++j) v[j] = c0 + j < count ? x[c0 + j] : T(0);
#pragma unroll
    for (uint32_t j = 0; j < kChunk; ++j) {
      if (c0 + j == 0)
        acc = v[0];
      else if (c0 + j < count)
        acc += v[j];
    }
  }
  return acc;
}

}  // namespace tg
