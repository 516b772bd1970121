// The mesh functional F = sum_c sum_q w_q det J + beta sum_c sum_q w_q |J|_F^2 / det(J)^(2/3)
// over a block of 3D cells, and its gradient with respect to the node coordinates (a vector-
// Jacobian product). The reference has no adjoint; the forward reuses the factor kernels'
// Jacobian arithmetic (tetraloma-geometry/src/layout.rs:138-213) and the reduction is
// deterministic like measure.rs:17-102 (fixed-order per-block trees, then one ordered pass).
//
// Forward: two lanes per cell (the even and the odd points, in two wavefronts); J in T (FP32:
// coordinates shifted to the cell's first node) at each point, the integrand in T, accumulated
// per lane in FP64; a fixed tree per workgroup writes one FP64 partial; a single-workgroup pass
// sums the partials in order. Backward: J is recomputed (not stored), dF/dJ is formed from it in
// T, each lane sums its points' contributions dF/dx(k, g) = sum_q sum_d dF/dJ(g, d)
// dphi_k/dxi_d, the two lanes of a cell are added (even + odd) and each (cell, local node)
// contribution is stored in T at its CSR position (node by node, ascending cell then local
// node); a node pass then sums each node's contiguous run in FP64 and scales by the seed, so the
// gradient is bitwise reproducible without atomics.
#include <hip/hip_runtime.h>

#include "adjoint_internal.hpp"

namespace tg {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ float tfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double tfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float tcbrt(float x) { return cbrtf(x); }
__device__ __forceinline__ double tcbrt(double x) { return cbrt(x); }

template <typename T>
__device__ __forceinline__ T det3(const T* j) {
  return j[0] * (j[4] * j[8] - j[7] * j[5]) - j[3] * (j[1] * j[8] - j[7] * j[2]) +
         j[6] * (j[1] * j[5] - j[4] * j[2]);
}

// Gathers the K nodes of cell c (indices clamped to the last node); SHIFT subtracts node 0.
template <typename T, int K, bool SHIFT>
__device__ __forceinline__ void gather(const T* __restrict__ nodes, const uint32_t* __restrict__ row,
                                       uint32_t last_node, T (&x)[K][3]) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
    uint32_t n = row[k];
    n = n < last_node ? n : last_node;
#pragma unroll
    for (int g = 0; g < 3; ++g) x[k][g] = nodes[static_cast<uint64_t>(n) * 3 + g];
  }
  if constexpr (SHIFT) {
#pragma unroll
    for (int k = 1; k < K; ++k)
#pragma unroll
      for (int g = 0; g < 3; ++g) x[k][g] -= x[0][g];
  }
}

// J (column-major, component d * 3 + g) of the gathered cell at point q.
template <typename T, int K, bool SHIFT>
__device__ __forceinline__ void jacobian(const T (&x)[K][3], const T* __restrict__ t, T* j) {
  constexpr int K0 = SHIFT ? 1 : 0;
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int g = 0; g < 3; ++g) {
      T acc = x[K0][g] * t[4 * K0 + 1 + d];
#pragma unroll
      for (int k = K0 + 1; k < K; ++k) acc = tfma(x[k][g], t[4 * k + 1 + d], acc);
      j[d * 3 + g] = acc;
    }
}

// Orders the LDS accesses of a wavefront before and after this point.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Fixed-order tree over the workgroup; thread 0 returns the sum.
__device__ __forceinline__ double block_sum(double v, double* lds) {
  lds[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(s)) lds[threadIdx.x] += lds[threadIdx.x + s];
    __syncthreads();
  }
  return lds[0];
}

// Both kernels give each cell two lanes in two wavefronts of the workgroup: wavefronts 2i and
// 2i + 1 hold the same 64 cells and take the even and the odd points, so the table addresses
// stay uniform per wavefront (scalar loads) and twice as many wavefronts share the work.
constexpr int kCells = kBlock / 2;  // cells per workgroup

template <typename T, int K, bool SHIFT>
__global__ void __launch_bounds__(kBlock)
    functional_forward(const T* __restrict__ nodes, const uint32_t* __restrict__ dofmap,
                       const T* __restrict__ tab, const double* __restrict__ wts, uint32_t nq,
                       uint64_t num_cells, uint32_t last_node, double beta,
                       double* __restrict__ partials) {
  __shared__ double lds[kBlock];
  // The wavefront index, made a scalar so that the table addresses are known to be uniform.
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), half = wave & 1u;
  const uint64_t c = static_cast<uint64_t>(blockIdx.x) * kCells + (wave >> 1) * 64 + (threadIdx.x & 63u);
  double cell_sum = 0.0;
  if (c < num_cells) {
    T x[K][3];
    gather<T, K, SHIFT>(nodes, dofmap + c * K, last_node, x);
    const T b = static_cast<T>(beta);
    for (uint32_t q = half; q < nq; q += 2) {
      T j[9];
      jacobian<T, K, SHIFT>(x, tab + static_cast<uint64_t>(q) * K * 4, j);
      const T det = det3(j);
      T n2 = j[0] * j[0];
#pragma unroll
      for (int i = 1; i < 9; ++i) n2 = tfma(j[i], j[i], n2);
      const T r = tcbrt(det);
      const T f = det + b * n2 / (r * r);
      cell_sum = tfma(wts[q], static_cast<double>(f), cell_sum);
    }
  }
  const double s = block_sum(cell_sum, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// Sums n partials in a fixed order into *out (one workgroup).
__global__ void __launch_bounds__(kBlock)
    sum_partials(const double* __restrict__ partials, uint32_t n, double* __restrict__ out) {
  __shared__ double lds[kBlock];
  double acc = 0.0;
  for (uint32_t i = threadIdx.x; i < n; i += kBlock) acc += partials[i];
  const double s = block_sum(acc, lds);
  if (threadIdx.x == 0) *out = s;
}

template <typename T, int K, bool SHIFT>
__global__ void __launch_bounds__(kBlock)
    functional_vjp_cells(const T* __restrict__ nodes, const uint32_t* __restrict__ dofmap,
                         const T* __restrict__ tab, const double* __restrict__ wts, uint32_t nq,
                         uint64_t num_cells, uint32_t last_node, double beta,
                         const uint32_t* __restrict__ slot, T* __restrict__ contrib) {
  __shared__ T lds[kBlock / 128][K * 3 * 64];  // per odd wavefront: all K x 3 sums of its 64 cells
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), half = wave & 1u,
                 w = threadIdx.x & 63u;
  const uint64_t c0 = static_cast<uint64_t>(blockIdx.x) * kCells + (wave >> 1) * 64;
  const uint64_t cl = c0 + w;
  const uint64_t c = cl < num_cells ? cl : num_cells - 1;
  const uint32_t valid =
      c0 >= num_cells ? 0u : (num_cells - c0 < 64 ? static_cast<uint32_t>(num_cells - c0) : 64u);
  T ct[K][3];
#pragma unroll
  for (int k = 0; k < K; ++k) ct[k][0] = ct[k][1] = ct[k][2] = T(0);
  if (valid > 0) {
    T x[K][3];
    gather<T, K, SHIFT>(nodes, dofmap + c * K, last_node, x);
    // Per-cell arithmetic in the geometry's type (J recomputed, dF/dJ, contributions summed over
    // the wavefront's points); the node reduction that follows is FP64.
    const T b = static_cast<T>(beta);
    for (uint32_t q = half; q < nq; q += 2) {
      const T* __restrict__ t = tab + static_cast<uint64_t>(q) * K * 4;
      T j[9];
      jacobian<T, K, SHIFT>(x, t, j);
      // Cofactors cof(g, d) = d det / d J(g, d), stored like J at d * 3 + g.
      T cof[9];
      cof[0] = j[4] * j[8] - j[7] * j[5];
      cof[3] = -(j[1] * j[8] - j[7] * j[2]);
      cof[6] = j[1] * j[5] - j[4] * j[2];
      cof[1] = -(j[3] * j[8] - j[6] * j[5]);
      cof[4] = j[0] * j[8] - j[6] * j[2];
      cof[7] = -(j[0] * j[5] - j[3] * j[2]);
      cof[2] = j[3] * j[7] - j[6] * j[4];
      cof[5] = -(j[0] * j[7] - j[6] * j[1]);
      cof[8] = j[0] * j[4] - j[3] * j[1];
      const T det = j[0] * cof[0] + j[3] * cof[3] + j[6] * cof[6];
      T n2 = T(0);
#pragma unroll
      for (int i = 0; i < 9; ++i) n2 = tfma(j[i], j[i], n2);
      const T r = tcbrt(det);
      const T inv23 = T(1) / (r * r);
      const T wq = static_cast<T>(wts[q]);
      const T a = wq * (T(1) - b * (T(2) / T(3)) * n2 * inv23 / det);
      const T bb = wq * b * T(2) * inv23;
      T G[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) G[i] = a * cof[i] + bb * j[i];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const T d0 = t[4 * k + 1], d1 = t[4 * k + 2], d2 = t[4 * k + 3];
#pragma unroll
        for (int g = 0; g < 3; ++g)
          ct[k][g] = tfma(G[g], d0, tfma(G[3 + g], d1, tfma(G[6 + g], d2, ct[k][g])));
      }
    }
  }
  // The cell's contributions (even points + odd points) go to their CSR positions: the odd
  // wavefront stages all of its sums in LDS, the even one adds and stores each triple.
  const uint32_t* __restrict__ slots = slot + c * K;
  T* __restrict__ pair = lds[wave >> 1];
  if (half) {
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int g = 0; g < 3; ++g) pair[(k * 3 + g) * 64 + w] = ct[k][g];
  }
  __syncthreads();
  if (!half && cl < num_cells) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      T* dst = contrib + static_cast<uint64_t>(slots[k]) * 3;
#pragma unroll
      for (int g = 0; g < 3; ++g) dst[g] = ct[k][g] + pair[(k * 3 + g) * 64 + w];
    }
  }
}

// grad[n] = seed * (sum of node n's contributions, in CSR order) in FP64: one lane per node,
// reading its contiguous run of contributions.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    csr_node_sum(const uint64_t* __restrict__ offsets, const T* __restrict__ contrib,
                 uint64_t num_nodes, const double* __restrict__ seed, double* __restrict__ grad) {
  const uint64_t n = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (n >= num_nodes) return;
  const uint64_t e0 = offsets[n], e1 = offsets[n + 1];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  // One entry is one vector load of its three components; four entries' loads are issued
  // before their (in-order) additions.
  struct Entry {
    T v[3];
  };
  const Entry* __restrict__ entries = reinterpret_cast<const Entry*>(contrib);
  for (uint64_t e = e0; e < e1; e += 4) {
    Entry v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = entries[e + i < e1 ? e + i : e0];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (e + i < e1) {
        a0 += static_cast<double>(v[i].v[0]);
        a1 += static_cast<double>(v[i].v[1]);
        a2 += static_cast<double>(v[i].v[2]);
      }
    }
  }
  const double s = seed ? *seed : 1.0;
  grad[n * 3] = s * a0;
  grad[n * 3 + 1] = s * a1;
  grad[n * 3 + 2] = s * a2;
}

__global__ void __launch_bounds__(kBlock) fill_zero(double* __restrict__ p, uint64_t n) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

template <typename T>
using FwdFn =void (*)(const T*, const uint32_t*, const T*, const double*, uint32_t, uint64_t,
                       uint32_t, double, double*);
template <typename T>
using VjpFn = void (*)(const T*, const uint32_t*, const T*, const double*, uint32_t, uint64_t,
                       uint32_t, double, const uint32_t*, T*);

// Register-resident 3D elements: P1-P3 tetrahedra, Q1-Q2 hexahedra, P1-P2 prisms, P1 pyramids.
#define TG_ADJ_SHAPES(X) X(4) X(10) X(20) X(8) X(27) X(6) X(18) X(5)

template <typename T>
FwdFn<T> fwd_kernel(uint32_t k) {
  constexpr bool kShift = sizeof(T) == 4;
#define TG_CASE(K_) \
  if (k == K_) return functional_forward<T, K_, kShift>;
  TG_ADJ_SHAPES(TG_CASE)
#undef TG_CASE
  return nullptr;
}

template <typename T>
VjpFn<T> vjp_kernel(uint32_t k) {
  constexpr bool kShift = sizeof(T) == 4;
#define TG_CASE(K_) \
  if (k == K_) return functional_vjp_cells<T, K_, kShift>;
  TG_ADJ_SHAPES(TG_CASE)
#undef TG_CASE
  return nullptr;
}

}  // namespace

uint32_t functional_partials(uint64_t num_cells) {
  return static_cast<uint32_t>((num_cells + kCells - 1) / kCells);
}

template <typename T>
hipError_t launch_functional_forward(const T* nodes, const uint32_t* dofmap, uint32_t k,
                                     const T* tab, const double* wts, uint32_t nq,
                                     uint64_t num_cells, uint32_t last_node, double beta,
                                     double* partials, double* out, hipStream_t stream) {
  FwdFn<T> fn = fwd_kernel<T>(k);
  if (!fn) return hipErrorNotSupported;
  const uint32_t blocks = functional_partials(num_cells);
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(kBlock), 0, stream, nodes, dofmap, tab, wts, nq,
                     num_cells, last_node, beta, partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(sum_partials, dim3(1), dim3(kBlock), 0, stream, partials, blocks, out);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_functional_vjp(const T* nodes, const uint32_t* dofmap, uint32_t k,
                                 const T* tab, const double* wts, uint32_t nq, uint64_t num_cells,
                                 uint32_t last_node, double beta, const double* seed, T* contrib,
                                 const uint64_t* offsets, const uint32_t* slots,
                                 uint64_t num_nodes, double* grad, hipStream_t stream) {
  VjpFn<T> fn = vjp_kernel<T>(k);
  if (!fn) return hipErrorNotSupported;
  const uint32_t blocks = functional_partials(num_cells);
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(kBlock), 0, stream, nodes, dofmap, tab, wts, nq,
                     num_cells, last_node, beta, slots, contrib);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const uint64_t nb = (num_nodes + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(csr_node_sum<T>, dim3(static_cast<uint32_t>(nb)), dim3(kBlock), 0, stream,
                     offsets, contrib, num_nodes, seed, grad);
  return hipGetLastError();
}

bool functional_supported(uint32_t k) { return fwd_kernel<float>(k) != nullptr; }

// Zeroes n doubles with a kernel (no memset node, so the call may sit in a captured graph).
hipError_t launch_fill_zero(double* p, uint64_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const uint64_t nb = (n + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(fill_zero, dim3(static_cast<uint32_t>(nb)), dim3(kBlock), 0, stream, p, n);
  return hipGetLastError();
}

template hipError_t launch_functional_forward<float>(const float*, const uint32_t*, uint32_t,
                                                     const float*, const double*, uint32_t,
                                                     uint64_t, uint32_t, double, double*,
                                                     double*, hipStream_t);
template hipError_t launch_functional_forward<double>(const double*, const uint32_t*, uint32_t,
                                                      const double*, const double*, uint32_t,
                                                      uint64_t, uint32_t, double, double*,
                                                      double*, hipStream_t);
template hipError_t launch_functional_vjp<float>(const float*, const uint32_t*, uint32_t,
                                                 const float*, const double*, uint32_t, uint64_t,
                                                 uint32_t, double, const double*, float*,
                                                 const uint64_t*, const uint32_t*, uint64_t,
                                                 double*, hipStream_t);
template hipError_t launch_functional_vjp<double>(const double*, const uint32_t*, uint32_t,
                                                  const double*, const double*, uint32_t, uint64_t,
                                                  uint32_t, double, const double*, double*,
                                                  const uint64_t*, const uint32_t*, uint64_t,
                                                  double*, hipStream_t);

}  // namespace tg
