// GLL -> equispaced conversion of the node array (tetraloma-geometry/src/equispaced.rs:1-12,
// equispaced_nodes): each node of a GLL block takes the value of its first occurrence in
// (block, cell, local node) order from the block's physical points at the equispaced lattice.
// The first occurrence of every node is the minimum of its flat occurrence indices, taken with
// atomicMin, which gives the same node-by-node answer in any execution order; each node is then
// written once, from that occurrence.
#include <hip/hip_runtime.h>

#include "equispaced_internal.hpp"

namespace tg {
namespace {

constexpr uint32_t kBlock = 256;

__global__ void __launch_bounds__(kBlock) fill_none(uint32_t* __restrict__ first, uint64_t n) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) first[i] = 0xffffffffu;
}

// first[node] = min over the block's occurrences i of start + i (nodes past the array skipped,
// as equispaced.rs:82 skips them).
__global__ void __launch_bounds__(kBlock)
    first_occurrence(const uint32_t* __restrict__ dofmap, uint64_t len, uint32_t start,
                     uint64_t num_nodes, uint32_t* __restrict__ first) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= len) return;
  const uint32_t node = dofmap[i];
  if (node < num_nodes) atomicMin(&first[node], start + static_cast<uint32_t>(i));
}

// out[node] = the field's value at the node's first occurrence, for the nodes whose first
// occurrence start + i (i < len) is in this block: occurrence i is local node i % k of cell
// c = i / k, read at lattice point point_of[i % k] of the field (tiles of 16). One thread per
// node, so the table is read in order and each node written once.
template <typename T, int GD>
__global__ void __launch_bounds__(kBlock)
    select_first(uint64_t len, uint32_t start, uint32_t k, const uint32_t* __restrict__ point_of,
                 const T* __restrict__ field, const uint32_t* __restrict__ first,
                 uint64_t num_nodes, T* __restrict__ out) {
  const uint64_t node = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (node >= num_nodes) return;
  const uint32_t o = first[node];  // 0xffffffff: no occurrence
  if (o < start || static_cast<uint64_t>(o - start) >= len) return;
  const uint64_t i = o - start;
  const uint64_t c = i / k, q = point_of[i - c * k];
#pragma unroll
  for (int g = 0; g < GD; ++g)
    out[node * GD + g] = field[(((c >> 4) * k + q) * GD + g) * 16 + (c & 15u)];
}

uint32_t grid_of(uint64_t n) { return static_cast<uint32_t>((n + kBlock - 1) / kBlock); }

}  // namespace

hipError_t launch_first_init(uint32_t* first, uint64_t num_nodes, hipStream_t stream) {
  if (num_nodes == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_none, dim3(grid_of(num_nodes)), dim3(kBlock), 0, stream, first,
                     num_nodes);
  return hipGetLastError();
}

hipError_t launch_first_occurrence(const uint32_t* dofmap, uint64_t len, uint32_t start,
                                   uint64_t num_nodes, uint32_t* first, hipStream_t stream) {
  if (len == 0) return hipSuccess;
  hipLaunchKernelGGL(first_occurrence, dim3(grid_of(len)), dim3(kBlock), 0, stream, dofmap, len,
                     start, num_nodes, first);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_select_first(int gdim, uint64_t len, uint32_t start, uint32_t k,
                               const uint32_t* point_of, const T* field, const uint32_t* first,
                               uint64_t num_nodes, T* out, hipStream_t stream) {
  if (len == 0 || num_nodes == 0) return hipSuccess;
  if ((num_nodes + kBlock - 1) / kBlock > 0x7fffffffu) return hipErrorInvalidConfiguration;
  const dim3 grid(grid_of(num_nodes));
  switch (gdim) {
    case 1:
      hipLaunchKernelGGL((select_first<T, 1>), grid, dim3(kBlock), 0, stream, len, start, k,
                         point_of, field, first, num_nodes, out);
      break;
    case 2:
      hipLaunchKernelGGL((select_first<T, 2>), grid, dim3(kBlock), 0, stream, len, start, k,
                         point_of, field, first, num_nodes, out);
      break;
    case 3:
      hipLaunchKernelGGL((select_first<T, 3>), grid, dim3(kBlock), 0, stream, len, start, k,
                         point_of, field, first, num_nodes, out);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template hipError_t launch_select_first<float>(int, uint64_t, uint32_t, uint32_t,
                                               const uint32_t*, const float*, const uint32_t*,
                                               uint64_t, float*, hipStream_t);
template hipError_t launch_select_first<double>(int, uint64_t, uint32_t, uint32_t,
                                                const uint32_t*, const double*, const uint32_t*,
                                                uint64_t, double*, hipStream_t);

}  // namespace tg
