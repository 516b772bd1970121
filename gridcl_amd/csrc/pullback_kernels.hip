// Pull-back of physical points to reference points by Newton iteration
// (tetraloma-geometry/src/pull_back.rs:26-180, pull_back_block :566).
//
// One wavefront per tile of 16 cells. The wavefront gathers the tile's nodes into LDS (indices
// clamped to the last node; FP32 coordinates shifted to each cell's node 0), forms each cell's
// coefficients in the element's spanning set, A(r, g) = sum_k C(r, k) x(k, g), into LDS, and then
// its lanes (16 cells x 4 points) run the Newton iteration of the cell's points: every point
// takes max_iterations steps from the reference cell's centroid, Delta = K(xi) (y - X(xi)) with K
// the pseudo-inverse of J, the new xi clamped to [-0.1, 1.1] per component, |Delta|_inf recorded
// as the last step, and the point frozen (xi and step kept) once that step is at most the
// tolerance. A point's result does not depend on its neighbours: the wavefront stops early only
// when all of its points are frozen, after which no step would change anything.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "jacobian.hpp"
#include "pullback_internal.hpp"
#include "spanning.hpp"

namespace tg {
namespace {

constexpr int kLanes = 64;

template <int CELL>
constexpr int cell_tdim() {
  return CELL == kInterval ? 1 : (CELL == kTriangle || CELL == kQuadrilateral) ? 2 : 3;
}

template <int CELL, int P>
constexpr int space_dim() {
  if constexpr (CELL == kInterval) return P + 1;
  if constexpr (CELL == kQuadrilateral) return (P + 1) * (P + 1);
  if constexpr (CELL == kHexahedron) return (P + 1) * (P + 1) * (P + 1);
  if constexpr (CELL == kTriangle) return (P + 1) * (P + 2) / 2;
  if constexpr (CELL == kTetrahedron) return (P + 1) * (P + 2) * (P + 3) / 6;
  if constexpr (CELL == kPrism) return (P + 1) * (P + 1) * (P + 2) / 2;
  return (P + 1) * (P + 2) * (2 * P + 3) / 6;  // pyramid
}

template <typename T>
__device__ __forceinline__ T tmax();
template <>
__device__ __forceinline__ float tmax<float>() {
  return FLT_MAX;
}
template <>
__device__ __forceinline__ double tmax<double>() {
  return DBL_MAX;
}

template <typename T, int CELL, int GD, int P>
__global__ void __launch_bounds__(kLanes)
    pull_back_tile(const T* __restrict__ nodes, const uint32_t* __restrict__ dofmap,
                   const T* __restrict__ coef, const T* __restrict__ x, T* __restrict__ xi_out,
                   T* __restrict__ step_out, PullBackParams<T> p) {
  constexpr int TD = cell_tdim<CELL>();
  constexpr int K = space_dim<CELL, P>();
  __shared__ T xs[K * GD * 16];  // [k][g][cell]
  __shared__ T as[K * GD * 16];  // [r][g][cell]
  const uint32_t lane = threadIdx.x, cell = lane & 15u;
  const uint64_t c0 = static_cast<uint64_t>(blockIdx.x) * 16;

  // Gather: item (k, cell) -> xs, cells past the block read the last cell (field.rs:234-253).
  for (uint32_t it = lane; it < 16u * K; it += kLanes) {
    const uint32_t cc = it & 15u, k = it >> 4;
    uint64_t c = c0 + cc;
    c = c < p.num_cells ? c : p.num_cells - 1;
    uint32_t n = dofmap[c * K + k];
    n = n < p.last_node ? n : p.last_node;
#pragma unroll
    for (int g = 0; g < GD; ++g) xs[(k * GD + g) * 16 + cc] = nodes[static_cast<uint64_t>(n) * GD + g];
  }
  __syncthreads();

  // Coefficients: lane (cell, g = lane / 16) forms A(r, g) for every r; C(r, k) is the same for
  // all lanes, so its loads are scalar.
  const uint32_t gl = lane >> 4;
  if (gl < static_cast<uint32_t>(GD)) {
    const T sh = p.shift ? xs[gl * 16 + cell] : T(0);
    for (int r = 0; r < K; ++r) {
      const T* __restrict__ cr = coef + r * K;
      T acc = cr[0] * (xs[gl * 16 + cell] - sh);
#pragma unroll 8
      for (int k = 1; k < K; ++k) acc += cr[k] * (xs[(k * GD + gl) * 16 + cell] - sh);
      as[(r * GD + gl) * 16 + cell] = acc;
    }
  }
  __syncthreads();

  const uint64_t c = c0 + cell;
  if (c >= p.padded_cells) return;  // past the field's last tile
  T sh[GD];
#pragma unroll
  for (int g = 0; g < GD; ++g) sh[g] = p.shift ? xs[g * 16 + cell] : T(0);
  const T* __restrict__ a = as + cell;
  const uint32_t tl = p.tile_log2, tmask = (1u << tl) - 1u;
  const uint64_t tile = c >> tl, in_tile = c & tmask;
  const T lo = static_cast<T>(-0.1f), hi = static_cast<T>(1.1f), big = tmax<T>();
  for (uint32_t q = lane >> 4; q < p.nq; q += 4) {
    T y[GD];
#pragma unroll
    for (int g = 0; g < GD; ++g)
      y[g] = (x[(((tile * p.nq + q) * GD + g) << tl) | in_tile] - p.origin[g]) - sh[g];
    T xi[TD];
#pragma unroll
    for (int d = 0; d < TD; ++d) xi[d] = p.centroid[d];
    T last = big;
    bool frozen = false;
    int offset = 0;
    for (uint32_t it = 0; it < p.max_iterations; ++it) {
      // Coefficients of more than 256 bytes per cell are read from LDS in every step: an offset
      // the compiler must treat as changed keeps it from holding them all in registers.
      if constexpr (K * GD * sizeof(T) > 256) asm volatile("" : "+v"(offset));
      T X[GD], J[TD * GD], Kp[TD * GD];
      span::coordinate_map<T, CELL, TD, GD, P>(a + offset, xi, X, J);
      jac::pinv_of<T, TD, GD>(J, Kp);
      T s = T(0);
      bool bad = false;
      T delta[TD];
#pragma unroll
      for (int t = 0; t < TD; ++t) {
        T acc = Kp[t] * (y[0] - X[0]);
#pragma unroll
        for (int g = 1; g < GD; ++g) acc += Kp[g * TD + t] * (y[g] - X[g]);
        delta[t] = acc;
        const T m = acc < T(0) ? -acc : acc;
        s = (t == 0 || m > s) ? m : s;
        bad = bad || !(m <= big);
      }
      if (bad) s = big;
      if (!frozen) {
#pragma unroll
        for (int d = 0; d < TD; ++d) {
          T v = xi[d] + delta[d];
          v = v < lo ? lo : v;
          v = v > hi ? hi : v;
          xi[d] = v;
        }
        last = s;
        frozen = s <= p.tol;
      }
      // Frozen points keep xi and step, so once every point of the wavefront is frozen the
      // remaining steps change nothing: leave them out (the results are those of all steps).
      if (__all(frozen)) break;
    }
#pragma unroll
    for (int d = 0; d < TD; ++d) xi_out[(((tile * p.nq + q) * TD + d) << tl) | in_tile] = xi[d];
    step_out[((tile * p.nq + q) << tl) | in_tile] = last;
  }
}

template <typename T>
using PullBackFn = void (*)(const T*, const uint32_t*, const T*, const T*, T*, T*,
                            PullBackParams<T>);

// Degrees 1-4; pyramids degree 1, hexahedra 1-3 (the degree-4 hexahedron's 125 nodes and
// coefficients per cell would take 96 KiB of LDS per 16 FP64 cells, above a workgroup's 64 KiB).
template <typename T, int CELL, int GD>
PullBackFn<T> by_degree(int degree) {
  if constexpr (CELL == kPyramid) {
    return degree == 1 ? pull_back_tile<T, CELL, GD, 1> : nullptr;
  } else {
    switch (degree) {
      case 1: return pull_back_tile<T, CELL, GD, 1>;
      case 2: return pull_back_tile<T, CELL, GD, 2>;
      case 3: return pull_back_tile<T, CELL, GD, 3>;
      case 4:
        if constexpr (CELL == kHexahedron) return nullptr;
        else return pull_back_tile<T, CELL, GD, 4>;
      default: return nullptr;
    }
  }
}

template <typename T>
PullBackFn<T> pick(int cell, int gdim, int degree) {
  switch (cell) {
    case kInterval:
      return gdim == 1 ? by_degree<T, kInterval, 1>(degree)
                       : gdim == 2 ? by_degree<T, kInterval, 2>(degree) : nullptr;
    case kTriangle:
      return gdim == 2 ? by_degree<T, kTriangle, 2>(degree)
                       : gdim == 3 ? by_degree<T, kTriangle, 3>(degree) : nullptr;
    case kQuadrilateral:
      return gdim == 2 ? by_degree<T, kQuadrilateral, 2>(degree)
                       : gdim == 3 ? by_degree<T, kQuadrilateral, 3>(degree) : nullptr;
    case kTetrahedron: return gdim == 3 ? by_degree<T, kTetrahedron, 3>(degree) : nullptr;
    case kHexahedron: return gdim == 3 ? by_degree<T, kHexahedron, 3>(degree) : nullptr;
    case kPrism: return gdim == 3 ? by_degree<T, kPrism, 3>(degree) : nullptr;
    case kPyramid: return gdim == 3 ? by_degree<T, kPyramid, 3>(degree) : nullptr;
    default: return nullptr;
  }
}

}  // namespace

bool pull_back_supported(int cell, int gdim, int degree) {
  return pick<float>(cell, gdim, degree) != nullptr;
}

template <typename T>
hipError_t launch_pull_back(int cell, int gdim, int degree, const T* nodes,
                            const uint32_t* dofmap, const T* coef, const T* x, T* xi, T* step,
                            const PullBackParams<T>& p, hipStream_t stream) {
  PullBackFn<T> fn = pick<T>(cell, gdim, degree);
  if (!fn) return hipErrorNotSupported;
  const uint64_t blocks = (p.padded_cells + 15) / 16;
  if (blocks == 0) return hipSuccess;
  if (blocks > 0x7fffffffu) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(fn, dim3(static_cast<uint32_t>(blocks)), dim3(kLanes), 0, stream, nodes,
                     dofmap, coef, x, xi, step, p);
  return hipGetLastError();
}

template hipError_t launch_pull_back<float>(int, int, int, const float*, const uint32_t*,
                                            const float*, const float*, float*, float*,
                                            const PullBackParams<float>&, hipStream_t);
template hipError_t launch_pull_back<double>(int, int, int, const double*, const uint32_t*,
                                             const double*, const double*, double*, double*,
                                             const PullBackParams<double>&, hipStream_t);

}  // namespace tg
