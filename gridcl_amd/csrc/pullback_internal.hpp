// The pull-back kernels (pullback_kernels.hip), for the C-ABI entry point in pullback.cpp.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tg {

// Scalars of one pull-back launch.
template <typename T>
struct PullBackParams {
  uint64_t num_cells;     // cells of the block (> 0)
  uint64_t padded_cells;  // num_cells rounded up to whole tiles of the fields
  uint32_t last_node;     // dofmap indices are clamped to it
  uint32_t nq;            // points per cell
  uint32_t tile_log2;     // log2 of the fields' tile
  uint32_t max_iterations;
  uint32_t shift;         // 1: coordinates relative to each cell's node 0 (FP32)
  T tol;                  // a point freezes once |Delta|_inf <= tol
  T origin[3];            // the geometry's origin narrowed to T (the points are absolute)
  T centroid[3];          // the start of the iteration
};

// Whether a kernel exists for (cell, gdim, degree).
bool pull_back_supported(int cell, int gdim, int degree);

// Launches the pull-back of block's points x (a field of gdim components) into xi (tdim
// components) and step (one); coef is the element's C [r][k] in the kernels' spanning order.
// Returns hipErrorNotSupported for an element without a kernel.
template <typename T>
hipError_t launch_pull_back(int cell, int gdim, int degree, const T* nodes,
                            const uint32_t* dofmap, const T* coef, const T* x, T* xi, T* step,
                            const PullBackParams<T>& p, hipStream_t stream);

}  // namespace tg
