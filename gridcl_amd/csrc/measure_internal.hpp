// The measure kernels (measure_kernels.hip), for the C-ABI entry point in measure.cpp.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

struct tg_geometry;
struct tg_pointset;

namespace tg {

// Elements of the geometry's type the measure of num_cells cells needs besides the det J field:
// the per-cell sums and the first blocked pass's outputs.
uint64_t measure_scratch_elems(uint64_t num_cells);

// *value (device) = the FP64 widening of the fixed-order sum over cells of sum_q det J w_q, from
// a det J field with tiles of 16 and nq points per cell; 0 for no cells or points. `scratch`
// holds measure_scratch_elems(num_cells) elements.
template <typename T>
hipError_t launch_measure(const T* detj, const double* weights, uint32_t nq, uint64_t num_cells,
                          T* scratch, double* value, hipStream_t stream);

// *value (device) = the FP64 widening of the passes of 64-value sums over the values
// scratch[0, len) (len > 0: the cell sums, or the first pass's outputs); scratch holds
// len + ceil(len / 64) elements and is overwritten.
template <typename T>
hipError_t launch_measure_sums(uint64_t len, T* scratch, double* value, hipStream_t stream);

// tg_measure_block with caller-provided work memory: `work` (256-byte aligned) of work_bytes
// bytes, or nullptr to allocate it stream-ordered for the call. With `work` nothing is
// allocated: the point set's device tables and weights must be resident (TG_ERR_INVALID_ARGUMENT
// otherwise), and a block that needs more than work_bytes fails with TG_ERR_TOO_LARGE; both
// before anything is enqueued.
int measure_block_with(const tg_geometry* g, int32_t block, tg_pointset* pts, uint32_t options,
                       double* value, void* work, size_t work_bytes, hipStream_t stream);

}  // namespace tg
