// Internal types shared by the C-ABI host code and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tg {

// Factor bits, identical to the TG_* factor enum of the public header.
enum : uint32_t {
  kFPoints = 1u << 0,
  kFJac = 1u << 1,
  kFDetJ = 1u << 2,
  kFPinv = 1u << 3,
  kFFacetNormals = 1u << 4,
  kFNormals = 1u << 5,
};

// Scalars of one factor launch. Device arrays are passed as separate restrict-qualified kernel
// parameters so that the table loads can stay in the scalar cache.
template <typename T>
struct FactorParams {
  uint64_t num_cells;     // cells of the block (> 0)
  uint64_t padded_cells;  // num_cells rounded up to whole tiles; lanes past num_cells copy the last
  uint32_t last_node;     // dofmap indices are clamped to it
  uint32_t nq;            // points
  uint32_t k;             // nodes per cell (used by the runtime-K kernel)
  uint32_t tile_log2;     // log2 of the QuadField tile
  uint32_t mask;          // factor bits
  uint32_t fixup;         // 1: recompute det J in FP64 where |det J| < tau |J|_F^3
  double tau;
  T origin[3];            // origin narrowed to T, added back to X
};

struct FactorOutputs {
  void* x;
  void* j;
  void* detj;
  void* pinv;
  void* fnormals;
  void* normals;
};

// Launches the factor kernel for (T, tdim, gdim, k); `shift` selects the per-cell origin shift,
// `generic` forces the runtime-K kernel. Returns hipSuccess or the launch error, and
// hipErrorNotSupported if no kernel exists for the shape.
template <typename T>
hipError_t launch_factor(int tdim, int gdim, bool shift, bool generic, const T* nodes,
                         const uint32_t* dofmap, const uint32_t* flags, const T* tab,
                         const double* tab64, const T* ref_normals, const FactorParams<T>& p,
                         const FactorOutputs& out, hipStream_t stream);

template <typename T>
hipError_t launch_evector(const T* nodes, const uint32_t* dofmap, uint64_t num_cells,
                          uint64_t padded, uint32_t k, int gdim, uint32_t last_node, T* out,
                          hipStream_t stream);

}  // namespace tg
