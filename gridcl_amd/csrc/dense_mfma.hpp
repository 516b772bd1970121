// The dense node sums on the matrix cores (dense_mfma.hip): X, J, det J, K and facet normals of
// 3D cells of 20 nodes and more, tiles of 16, at most kDenseMfmaMaxPoints points.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "geom_internal.hpp"

namespace tg {

// taba is the table in operand order, [ceil(nq / 16)][4 kinds][ceil(k / 4)][64 lanes]: lane l of
// block qb, kind v, step s holds T_v[16 qb + l % 16][4 s + l / 16] (0 outside nq and k), with
// T_0 = phi and T_d = d phi / d xi_d, rounded once from FP64 (FP32 geometries) or in FP64.
// ref_normals: the point set's reference facet normals (needed for facet normals only).
// Each launch returns hipErrorNotSupported for the shapes it does not take.
constexpr uint32_t kDenseMfmaMaxPoints = 128;
size_t dense_mfma_table_elems(uint32_t k, uint32_t nq);

// FP32 geometries with the per-cell shift: 3D cells of 20 to 84 nodes in registers, more in the
// runtime-k form.
bool has_dense_mfma(int tdim, int gdim, uint32_t k);
hipError_t launch_dense_mfma(uint32_t k, const float* nodes, const uint32_t* dofmap,
                             const float* taba, const double* tab64, const float* ref_normals,
                             const FactorParams<float>& p, const FactorOutputs& out,
                             hipStream_t stream);
// The same on FP64-stored nodes: each offset from node 0 taken in FP64 and rounded to FP32 (as
// the register kernels on FP64 nodes take it), the FP64 recomputation from the FP64 nodes.
hipError_t launch_dense_mfma_f64_nodes(uint32_t k, const double* nodes, const uint32_t* dofmap,
                                       const float* taba, const double* tab64,
                                       const float* ref_normals, const FactorParams<float>& p,
                                       const FactorOutputs& out, hipStream_t stream);
// FP64 geometries (v_mfma_f64_16x16x4f64): 3D cells of k >= 20 nodes, the table in the same
// operand order in FP64, sums of the absolute coordinates from node 0 as the FP64 kernels take
// them.
bool has_dense_mfma_fp64(int tdim, int gdim, uint32_t k);
hipError_t launch_dense_mfma_fp64(uint32_t k, const double* nodes, const uint32_t* dofmap,
                                  const double* taba, const double* tab64,
                                  const double* ref_normals, const FactorParams<double>& p,
                                  const FactorOutputs& out, hipStream_t stream);

}  // namespace tg
