// tg_measure_block_noalloc: the block measure with its work memory in a fixed device array of
// the library instead of a stream-ordered allocation, for callers that cannot allocate (code
// under graph capture, serving loops, harnesses that forbid hipMalloc). It enqueues exactly the
// kernels of tg_measure_block (measure.cpp) and allocates nothing: not the work memory, and not
// the point set's device tables, which must be resident already.
//
// The array lives in every device's copy of the code object; its device address is taken once
// per device with hipGetSymbolAddress and kept, and only that address is passed to kernels. All
// calls on one device share the array. Nothing in it needs to start at zero: every value the
// kernels read was written earlier in the same call. Calls that overlap on one device race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/tetraloma_geom.h"
#include "handles.hpp"
#include "measure_internal.hpp"

namespace {

constexpr size_t kScratchBytes = TG_MEASURE_SCRATCH_BYTES;
static_assert(kScratchBytes % 256 == 0, "the work array is a whole number of 256-byte lines");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch`, per device

// The device address of `scratch` on the current device, looked up on its first use there.
int scratch_address(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return tg::hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices)
    return tg::fail(TG_ERR_INVALID_ARGUMENT, "device index " + std::to_string(dev) +
                                                 " beyond the " + std::to_string(kMaxDevices) +
                                                 " devices the work arrays are kept for");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return tg::hip_fail(e, "hipGetSymbolAddress(measure work array)");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return TG_OK;
}

}  // namespace

extern "C" {

int tg_measure_block_noalloc(const tg_geometry* g, int32_t block, tg_pointset* pts,
                             uint32_t options, double* value, void* stream) {
  void* work = nullptr;
  const int rc = scratch_address(&work);
  if (rc != TG_OK) return rc;
  return tg::measure_block_with(g, block, pts, options, value, work, kScratchBytes,
                                static_cast<hipStream_t>(stream));
}

}  // extern "C"
