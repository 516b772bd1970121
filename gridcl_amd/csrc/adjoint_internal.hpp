// Launchers of the functional and its vector-Jacobian product (adjoint_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tg {

// Number of FP64 partial sums the forward writes for num_cells cells.
uint32_t functional_partials(uint64_t num_cells);
// Whether a register-resident functional kernel exists for k nodes per cell.
bool functional_supported(uint32_t k);
// Zeroes n doubles with a kernel.
hipError_t launch_fill_zero(double* p, uint64_t n, hipStream_t stream);

template <typename T>
hipError_t launch_functional_forward(const T* nodes, const uint32_t* dofmap, uint32_t k,
                                     const T* tab, const double* wts, uint32_t nq,
                                     uint64_t num_cells, uint32_t last_node, double beta,
                                     double* partials, double* out, hipStream_t stream);

// contrib: [cells * k][3] scratch of the geometry's type in CSR order; slots[cell * k + k'] is
// the CSR position of (cell, local node k'); offsets [nodes + 1] delimit each node's run.
template <typename T>
hipError_t launch_functional_vjp(const T* nodes, const uint32_t* dofmap, uint32_t k,
                                 const T* tab, const double* wts, uint32_t nq, uint64_t num_cells,
                                 uint32_t last_node, double beta, const double* seed, T* contrib,
                                 const uint64_t* offsets, const uint32_t* slots,
                                 uint64_t num_nodes, double* grad, hipStream_t stream);

}  // namespace tg
