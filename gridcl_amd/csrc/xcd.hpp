// Workgroup-to-work remapping for the eight XCDs of an MI355X (each with its own L2).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tg {

// Hardware dispatch places workgroup b on XCD b % 8. Give XCD x the contiguous range of logical
// blocks [x * per + min(x, rem), ...): a bijection of [0, nb).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
  const uint32_t x = b & 7u, i = b >> 3;
  const uint32_t per = nb >> 3, rem = nb & 7u;
  return x * per + (x < rem ? x : rem) + i;
}

}  // namespace tg
