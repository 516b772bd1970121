// The GLL -> equispaced conversion kernels (equispaced_kernels.hip), for equispaced.cpp.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tg {

// first[0..num_nodes) = no occurrence yet.
hipError_t launch_first_init(uint32_t* first, uint64_t num_nodes, hipStream_t stream);
// Records the block's occurrences start + i (i < len, node dofmap[i]) in first.
hipError_t launch_first_occurrence(const uint32_t* dofmap, uint64_t len, uint32_t start,
                                   uint64_t num_nodes, uint32_t* first, hipStream_t stream);
// Writes out[node] from the field (X at the block's lattice, tiles of 16, gdim components; local
// node j at lattice point point_of[j]) for the nodes whose first occurrence is one of this
// block's, start .. start + len - 1.
template <typename T>
hipError_t launch_select_first(int gdim, uint64_t len, uint32_t start, uint32_t k,
                               const uint32_t* point_of, const T* field, const uint32_t* first,
                               uint64_t num_nodes, T* out, hipStream_t stream);

}  // namespace tg
