// The measure of a block, sum_c sum_q w_q det J (tetraloma-geometry/src/measure.rs:17-102, 141),
// as fixed-order sums on the device: each cell's sum over its points, then passes of blocked
// sums of 64 values, all in the geometry's type, and the total widened to FP64.
#include <hip/hip_runtime.h>

#include "measure_internal.hpp"
#include "measure_sums.hpp"

namespace tg {
namespace {

constexpr uint32_t kBlock = 256;

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
  const uint64_t count = start + kSumBlock < len ? kSumBlock : len - start;
  out[i] = ordered_sum<T>(x + start, static_cast<uint32_t>(count));
}

constexpr uint32_t kFinish = 1024;  // threads of the finishing workgroup

// The passes of 64-value sums from len <= 64 kFinish values x down to one (measure.rs:46-72),
// in one workgroup with the levels after the first in LDS, and the total widened to FP64
// (measure.rs:75-102). Every thread runs the same number of passes.
template <typename T>
__global__ void __launch_bounds__(kFinish)
    finish_sums(const T* __restrict__ x, uint64_t len, double* __restrict__ value) {
  __shared__ T level[2][kFinish];
  uint32_t n = static_cast<uint32_t>((len + kSumBlock - 1) / kSumBlock);
  for (uint32_t i = threadIdx.x; i < n; i += kFinish) {
    const uint64_t start = static_cast<uint64_t>(i) * kSumBlock;
    const uint64_t count = start + kSumBlock < len ? kSumBlock : len - start;
    level[0][i] = ordered_sum<T>(x + start, static_cast<uint32_t>(count));
  }
  __syncthreads();
  uint32_t cur = 0;
  while (n > 1) {
    const uint32_t m = (n + kSumBlock - 1) / kSumBlock;
    for (uint32_t i = threadIdx.x; i < m; i += kFinish) {
      const uint32_t start = i * kSumBlock;
      const uint32_t count = start + kSumBlock < n ? kSumBlock : n - start;
      level[cur ^ 1u][i] = ordered_sum<T>(&level[cur][start], count);
    }
    __syncthreads();
    cur ^= 1u;
    n = m;
  }
  if (threadIdx.x == 0) *value = static_cast<double>(level[cur][0]);
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
hipError_t launch_measure_sums(uint64_t len, T* scratch, double* value, hipStream_t stream) {
  if (len == 0 || (len + kBlock - 1) / kBlock > 0x7fffffffu) return hipErrorInvalidConfiguration;
  // Passes over the whole grid while one workgroup cannot finish; they ping-pong between the
  // values [len] and the space behind them.
  T* a = scratch;
  T* b = scratch + len;
  hipError_t e = hipSuccess;
  while (e == hipSuccess && len > static_cast<uint64_t>(kSumBlock) * kFinish) {
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
  hipLaunchKernelGGL(finish_sums<T>, dim3(1), dim3(kFinish), 0, stream, a, len, value);
  return hipGetLastError();
}

template hipError_t launch_measure<float>(const float*, const double*, uint32_t, uint64_t, float*,
                                          double*, hipStream_t);
template hipError_t launch_measure<double>(const double*, const double*, uint32_t, uint64_t,
                                           double*, double*, hipStream_t);
template hipError_t launch_measure_sums<float>(uint64_t, float*, double*, hipStream_t);
template hipError_t launch_measure_sums<double>(uint64_t, double*, double*, hipStream_t);

}  // namespace tg
