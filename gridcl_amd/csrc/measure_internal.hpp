// The measure kernels (measure_kernels.hip), for the C-ABI entry point in measure.cpp.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

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

}  // namespace tg
