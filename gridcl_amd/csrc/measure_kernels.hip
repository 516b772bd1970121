// The measure of a block, sum_c sum_q w_q det J (tetraloma-geometry/src/measure.rs:17-102, 141),
// as fixed-order sums on the device: each cell's sum over its points, then passes of blocked
// sums of 64 values, all in the geometry's type, and the total widened to FP64.
#include <hip/hip_runtime.h>

#include "measure_internal.hpp"

namespace tg {
namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kSumBlock = 64;  // values per output of one blocked pass (measure.rs:44)

// out[c] = sum_q det J(c, q) w_q for the cells c < num_cells of a det J field tiled by 16
// (cell c at ((c / 16) nq + q) 16 + c % 16), starting from q = 0 (measure.rs:17-41); the
// weights are narrowed to T as the reference's device weights are. Each term is rounded before
// it is added, so the sum can be repeated exactly on a host.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    cell_sums(const T* __restrict__ detj, const double* __restrict__ weights, uint32_t nq,
              uint64_t num_cells, T* __restrict__ out) {
  // A rounded product, then a rounded sum, as the expression reads (no fused multiply-add).
#pragma clang fp contract(off)
  const uint64_t c = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (c >= num_cells) return;
  const T* __restrict__ d = detj + (c >> 4) * nq * 16 + (c & 15u);
  T acc = d[0] * static_cast<T>(weights[0]);
  for (uint32_t q = 1; q < nq; ++q) acc += d[q * 16] * static_cast<T>(weights[q]);
  out[c] = acc;
}

// out[i] = x[64 i] + x[64 i + 1] + ... over the entries below len, in order (measure.rs:46-72).
template <typename T>
__global__ void __launch_bounds__(kBlock)
    block_sums(const T* __restrict__ x, uint64_t len, uint64_t num_out, T* __restrict__ out) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= num_out) return;
  const uint64_t start = i * kSumBlock;
  const uint64_t end = start + kSumBlock < len ? start + kSumBlock : len;
  T acc = x[start];
  for (uint64_t j = start + 1; j < end; ++j) acc += x[j];
  out[i] = acc;
}

template <typename T>
__global__ void widen(const T* __restrict__ x, double* __restrict__ value) {
  if (threadIdx.x == 0) *value = static_cast<double>(x[0]);
}

__global__ void store_zero(double* __restrict__ value) {
  if (threadIdx.x == 0) *value = 0.0;
}

uint32_t grid_of(uint64_t n) { return static_cast<uint32_t>((n + kBlock - 1) / kBlock); }

}  // namespace

uint64_t measure_scratch_elems(uint64_t num_cells) {
  return num_cells + (num_cells + kSumBlock - 1) / kSumBlock;
}

template <typename T>
hipError_t launch_measure(const T* detj, const double* weights, uint32_t nq, uint64_t num_cells,
                          T* scratch, double* value, hipStream_t stream) {
  if (num_cells == 0 || nq == 0) {
    hipLaunchKernelGGL(store_zero, dim3(1), dim3(64), 0, stream, value);
    return hipGetLastError();
  }
  if ((num_cells + kBlock - 1) / kBlock > 0x7fffffffu) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(cell_sums<T>, dim3(grid_of(num_cells)), dim3(kBlock), 0, stream, detj,
                     weights, nq, num_cells, scratch);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_measure_sums<T>(num_cells, scratch, value, stream);
}

template <typename T>
hipError_t launch_measure_sums(uint64_t num_cells, T* scratch, double* value, hipStream_t stream) {
  if (num_cells == 0 || (num_cells + kBlock - 1) / kBlock > 0x7fffffffu)
    return hipErrorInvalidConfiguration;
  // Ping-pong between the cell sums [num_cells] and the first pass's outputs behind them.
  T* a = scratch;
  T* b = scratch + num_cells;
  hipError_t e = hipSuccess;
  uint64_t len = num_cells;
  while (e == hipSuccess && len > 1) {
    const uint64_t num_out = (len + kSumBlock - 1) / kSumBlock;
    hipLaunchKernelGGL(block_sums<T>, dim3(grid_of(num_out)), dim3(kBlock), 0, stream, a, len,
                       num_out, b);
    e = hipGetLastError();
    T* t = a;
    a = b;
    b = t;
    len = num_out;
  }
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(widen<T>, dim3(1), dim3(64), 0, stream, a, value);
  return hipGetLastError();
}

template hipError_t launch_measure<float>(const float*, const double*, uint32_t, uint64_t, float*,
                                          double*, hipStream_t);
template hipError_t launch_measure<double>(const double*, const double*, uint32_t, uint64_t,
                                           double*, double*, hipStream_t);
template hipError_t launch_measure_sums<float>(uint64_t, float*, double*, hipStream_t);
template hipError_t launch_measure_sums<double>(uint64_t, double*, double*, hipStream_t);

}  // namespace tg
