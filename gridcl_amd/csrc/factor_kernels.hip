// Factor kernels for gfx950: X, J, det J, K, facet normals and normals of one cell block at a
// point set, all requested factors in one fused pass.
//
// Arithmetic follows the reference's factor kernel (tetraloma-geometry/src/layout.rs:138-369):
// J(g, d) = sum_k x(k, g) dphi_k/dxi_d, component d * gdim + g, sums starting from the first
// product; det J by first-row cofactors; the pseudo-determinant, pseudo-inverse, facet normals
// and normals by the same formulas. The layout of each output is the QuadField tile layout
// (field.rs:126-134).
//
// Mapping: one lane per cell. The lane gathers its cell's node coordinates once through the
// dofmap into registers (FP32: shifted to the cell's first node, so that the contraction works on
// offsets of the cell's own size); the basis tables are indexed uniformly across the wavefront,
// so they are read through the scalar cache; stores are non-temporal. With tiles of 16 cells a
// wavefront writes four 64-byte runs per component, and consecutive components fill the rest of
// each tile row. Blocks are remapped so that each XCD walks one contiguous cell range, keeping
// the nodes shared by neighbouring cells in that XCD's L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "geom_internal.hpp"
#include "jacobian.hpp"
#include "xcd.hpp"

namespace tg {
namespace {

using jac::det_of;
using jac::pinv_of;

constexpr int kBlock = 256;

__device__ __forceinline__ float tsqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double tsqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float tfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double tfma(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Global-memory views of pointers. The kernels' outputs, nodes, dofmaps and tables all live in
// global memory; accessed through generic pointers read from memory (the mixed-mesh kernel's
// block records) they would become flat accesses, which also count against the LDS wait counter,
// so that every wait for an LDS read would wait for the stores in flight as well.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> gp(T* p) {
  return (gptr<T>)p;
}

template <typename T>
__device__ __forceinline__ void st(void* base, uint64_t idx, T v) {
  __builtin_nontemporal_store(v, gp(reinterpret_cast<T*>(base)) + idx);
}

// Stores every factor of J that the mask asks for at point q (layout.rs:275-369).
template <typename T, int TD, int GD>
__device__ __forceinline__ void store_from_j(const T* j, T detj, const T* __restrict__ ref_normals,
                                             const uint32_t* __restrict__ flags, uint64_t c,
                                             uint32_t q, uint32_t mask, uint64_t row,
                                             uint32_t tl, uint32_t lane, const FactorOutputs& o) {
#pragma clang fp contract(off)  // as in jacobian.hpp: the same values in every kernel
  constexpr int JL = TD * GD;
  if (mask & kFJac) {
#pragma unroll
    for (int i = 0; i < JL; ++i) st<T>(o.j, (((row * JL) + i) << tl) | lane, j[i]);
  }
  if (mask & kFDetJ) st<T>(o.detj, (row << tl) | lane, detj);
  if (mask & (kFPinv | kFFacetNormals)) {
    T k[JL];
    pinv_of<T, TD, GD>(j, k);
    if (mask & kFPinv) {
#pragma unroll
      for (int i = 0; i < JL; ++i) st<T>(o.pinv, (((row * JL) + i) << tl) | lane, k[i]);
    }
    if constexpr (TD == GD) {
      if (mask & kFFacetNormals) {
        T m[GD];
        T s2 = T(0);
#pragma unroll
        for (int g = 0; g < GD; ++g) {
          T acc = k[g * TD] * gp(ref_normals)[q * TD];
#pragma unroll
          for (int t = 1; t < TD; ++t) acc += k[g * TD + t] * gp(ref_normals)[q * TD + t];
          m[g] = acc;
          s2 = g == 0 ? acc * acc : s2 + acc * acc;
        }
        const T nrm = tsqrt(s2);
#pragma unroll
        for (int g = 0; g < GD; ++g) st<T>(o.fnormals, (((row * GD) + g) << tl) | lane, m[g] / nrm);
      }
    }
  }
  if constexpr (TD + 1 == GD) {
    if (mask & kFNormals) {
      T n[GD];
      if constexpr (GD == 2) {
        n[0] = j[1];
        n[1] = -j[0];
      } else {
        n[0] = j[1] * j[5] - j[2] * j[4];
        n[1] = j[2] * j[3] - j[0] * j[5];
        n[2] = j[0] * j[4] - j[1] * j[3];
      }
      T s2 = n[0] * n[0];
#pragma unroll
      for (int g = 1; g < GD; ++g) s2 += n[g] * n[g];
      const T nrm = tsqrt(s2);
      const T s = (flags != nullptr && (gp(flags)[c] & 1u)) ? T(-1) : T(1);
#pragma unroll
      for (int g = 0; g < GD; ++g) st<T>(o.normals, (((row * GD) + g) << tl) | lane, (s * n[g]) / nrm);
    }
  }
}

// FP64 det J of cell c at point q from the stored coordinates (of type TN), for near-degenerate
// points.
template <int TD, int GD, typename TN>
__device__ __noinline__ double detj_fp64(const TN* __restrict__ nodes,
                                         const uint32_t* __restrict__ row, uint32_t k,
                                         uint32_t last_node, const double* __restrict__ tab64,
                                         uint32_t q) {
  double x0[GD];
  uint32_t n0 = row[0] < last_node ? row[0] : last_node;
#pragma unroll
  for (int g = 0; g < GD; ++g) x0[g] = static_cast<double>(nodes[static_cast<uint64_t>(n0) * GD + g]);
  double j[TD * GD];
#pragma unroll
  for (int i = 0; i < TD * GD; ++i) j[i] = 0.0;
  for (uint32_t kk = 1; kk < k; ++kk) {
    const uint32_t n = row[kk] < last_node ? row[kk] : last_node;
    const double* t = tab64 + (static_cast<uint64_t>(q) * k + kk) * 4;
#pragma unroll
    for (int g = 0; g < GD; ++g) {
      const double d = static_cast<double>(nodes[static_cast<uint64_t>(n) * GD + g]) - x0[g];
#pragma unroll
      for (int dd = 0; dd < TD; ++dd) j[dd * GD + g] = __builtin_fma(d, t[1 + dd], j[dd * GD + g]);
    }
  }
  return det_of<double, TD, GD>(j);
}

// Orders the LDS accesses of a wavefront before and after this point.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct Vec16<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};

// Gathers the K nodes of cell c (indices clamped to the last node); with SHIFT, nodes 1..K-1
// become offsets from node 0. Nodes stored in FP64 for FP32 arithmetic (TN = double, T = float,
// SHIFT): the offsets are taken in FP64 and rounded once, node 0 is rounded to FP32.
template <typename T, int GD, int K, bool SHIFT, typename TN = T>
__device__ __forceinline__ void gather_cell(const TN* __restrict__ nodes,
                                            const uint32_t* __restrict__ row, uint32_t last_node,
                                            T (&x)[K][GD]) {
  if constexpr (sizeof(TN) != sizeof(T)) {
    static_assert(SHIFT && sizeof(TN) > sizeof(T), "wider node storage needs the cell shift");
    TN x0[GD];
    const uint32_t n0 = row[0] < last_node ? row[0] : last_node;
#pragma unroll
    for (int g = 0; g < GD; ++g) {
      x0[g] = nodes[static_cast<uint64_t>(n0) * GD + g];
      x[0][g] = static_cast<T>(x0[g]);
    }
#pragma unroll
    for (int k = 1; k < K; ++k) {
      uint32_t n = row[k];
      n = n < last_node ? n : last_node;
#pragma unroll
      for (int g = 0; g < GD; ++g)
        x[k][g] = static_cast<T>(nodes[static_cast<uint64_t>(n) * GD + g] - x0[g]);
    }
    return;
  } else {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      uint32_t n = row[k];
      n = n < last_node ? n : last_node;
#pragma unroll
      for (int g = 0; g < GD; ++g) x[k][g] = nodes[static_cast<uint64_t>(n) * GD + g];
    }
    if constexpr (SHIFT) {
#pragma unroll
      for (int k = 1; k < K; ++k)
#pragma unroll
        for (int g = 0; g < GD; ++g) x[k][g] -= x[0][g];
    }
  }
}

// X (without the origin) and J of the gathered cell at the point whose table row is t; FP32
// adds node 0 back to X after the contraction of the offsets.
// The (value, d/dxi_0, d/dxi_1, d/dxi_2) row of node k at one point: one 16-byte load in FP32,
// two in FP64.
template <typename T>
struct TabRow {
  T v[4];
};
template <typename T>
__device__ __forceinline__ TabRow<T> load_row(const T* __restrict__ t) {
  using V = typename Vec16<T>::type;
  TabRow<T> r;
  if constexpr (sizeof(T) == 4) {
    const V a = *reinterpret_cast<const V*>(t);
    r.v[0] = a.x;
    r.v[1] = a.y;
    r.v[2] = a.z;
    r.v[3] = a.w;
  } else {
    const V a = *reinterpret_cast<const V*>(t);
    const V b = *reinterpret_cast<const V*>(t + 2);
    r.v[0] = a.x;
    r.v[1] = a.y;
    r.v[2] = b.x;
    r.v[3] = b.y;
  }
  return r;
}

// X (without the origin) and J of the gathered cell at the point whose table rows start at t
// (16-byte aligned); FP32 adds node 0 back to X after the contraction of the offsets. Each sum
// starts from its first product and runs in ascending node order.
template <typename T, int TD, int GD, int K, bool SHIFT, bool WX, bool WJ>
__device__ __forceinline__ void contract_fixed(const T (&x)[K][GD], const T* __restrict__ t, T* xv,
                                               T* j) {
  constexpr int K0 = SHIFT ? 1 : 0;  // first node whose offset is not identically zero
  T ax[GD], aj[TD * GD];
#pragma unroll
  for (int k = K0; k < K; ++k) {
    const TabRow<T> r = load_row<T>(t + 4 * k);
#pragma unroll
    for (int g = 0; g < GD; ++g) {
      if constexpr (WX) ax[g] = k == K0 ? x[k][g] * r.v[0] : tfma(x[k][g], r.v[0], ax[g]);
      if constexpr (WJ) {
#pragma unroll
        for (int d = 0; d < TD; ++d)
          aj[d * GD + g] =
              k == K0 ? x[k][g] * r.v[1 + d] : tfma(x[k][g], r.v[1 + d], aj[d * GD + g]);
      }
    }
  }
  if constexpr (WX) {
#pragma unroll
    for (int g = 0; g < GD; ++g) xv[g] = SHIFT ? x[0][g] + ax[g] : ax[g];
  }
  if constexpr (WJ) {
#pragma unroll
    for (int i = 0; i < TD * GD; ++i) j[i] = aj[i];
  }
}

// The factors asked for pick one straight-line contraction (a branch on uniform flags rather
// than a select per product).
template <typename T, int TD, int GD, int K, bool SHIFT>
__device__ __forceinline__ void contract(const T (&x)[K][GD], const T* __restrict__ t,
                                         bool want_x, bool want_j, T* xv, T* j) {
  if (want_x && want_j)
    contract_fixed<T, TD, GD, K, SHIFT, true, true>(x, t, xv, j);
  else if (want_j)
    contract_fixed<T, TD, GD, K, SHIFT, false, true>(x, t, xv, j);
  else
    contract_fixed<T, TD, GD, K, SHIFT, true, false>(x, t, xv, j);
}

// det J, recomputed in FP64 from the stored coordinates where |det J| < tau |J|_F^3 (FP32), for
// a cell of k nodes.
template <typename T, int TD, int GD, typename TN>
__device__ __forceinline__ T detj_checked_k(const T* j, const FactorParams<T>& p,
                                            const TN* __restrict__ nodes,
                                            const uint32_t* __restrict__ row, uint32_t k,
                                            const double* __restrict__ tab64, uint32_t q) {
  T detj = det_of<T, TD, GD>(j);
  if constexpr (TD == GD && sizeof(T) == 4) {
    if (p.fixup) {
      T n2 = T(0);
#pragma unroll
      for (int i = 0; i < TD * GD; ++i) n2 = tfma(j[i], j[i], n2);
      if (__builtin_fabsf(detj) < static_cast<T>(p.tau) * n2 * tsqrt(n2))
        detj = static_cast<T>(detj_fp64<TD, GD>(nodes, row, k, p.last_node, tab64, q));
    }
  }
  return detj;
}

template <typename T, int TD, int GD, int K, typename TN>
__device__ __forceinline__ T detj_checked(const T* j, const FactorParams<T>& p,
                                          const TN* __restrict__ nodes,
                                          const uint32_t* __restrict__ row,
                                          const double* __restrict__ tab64, uint32_t q) {
  return detj_checked_k<T, TD, GD, TN>(j, p, nodes, row, K, tab64, q);
}

// All requested factors of the cell of lane cl (cl < padded cells; lanes past the last cell
// compute the last cell, the padding of the field), stored straight from registers. Any tile.
template <typename T, int TD, int GD, int K, bool SHIFT, typename TN = T>
__device__ __forceinline__ void direct_cell(const TN* __restrict__ nodes,
                                            const uint32_t* __restrict__ dofmap,
                                            const uint32_t* __restrict__ flags,
                                            const T* __restrict__ tab,
                                            const double* __restrict__ tab64,
                                            const T* __restrict__ ref_normals,
                                            const FactorParams<T>& p, const FactorOutputs& o,
                                            uint64_t cl) {
  constexpr int JL = TD * GD;
  const uint64_t c = cl < p.num_cells ? cl : p.num_cells - 1;
  const uint32_t* __restrict__ row = dofmap + c * K;
  T x[K][GD];
  gather_cell<T, GD, K, SHIFT, TN>(nodes, row, p.last_node, x);
  const uint32_t tl = p.tile_log2;
  const uint64_t tile = cl >> tl;
  const uint32_t lane = static_cast<uint32_t>(cl) & ((1u << tl) - 1u);
  const uint32_t mask = p.mask;
  const bool need_j = (mask & ~kFPoints) != 0;
  for (uint32_t q = 0; q < p.nq; ++q) {
    const uint64_t row_q = tile * p.nq + q;  // (tile * nq + q): times LEN and the component next
    T xv[GD], j[JL];
    contract<T, TD, GD, K, SHIFT>(x, tab + static_cast<uint64_t>(q) * K * 4,
                                  (mask & kFPoints) != 0, need_j, xv, j);
    if (mask & kFPoints) {
#pragma unroll
      for (int g = 0; g < GD; ++g) st<T>(o.x, (((row_q * GD) + g) << tl) | lane, xv[g] + p.origin[g]);
    }
    if (need_j) {
      const T detj = (mask & kFDetJ) ? detj_checked<T, TD, GD, K>(j, p, nodes, row, tab64, q) : T(0);
      store_from_j<T, TD, GD>(j, detj, ref_normals, flags, c, q, mask, row_q, tl, lane, o);
    }
  }
}

// Copies one staged field of a wavefront's four tiles at point q: for each tile the LEN * 16
// values of (tile, q) are contiguous in the staging buffer and in the QuadField. Tiles at or past
// `ntiles` are padding of the launch and skipped.
template <typename T, int LEN>
__device__ __forceinline__ void copy_point(const T* __restrict__ stage, void* out, uint64_t tile0,
                                           uint64_t ntiles, uint32_t nq, uint32_t q, uint32_t w) {
  using V = typename Vec16<T>::type;
  constexpr uint32_t E = 16 / sizeof(T);
  constexpr uint32_t per_tile = LEN * 16 / E;
  for (uint32_t s = w; s < 4 * per_tile; s += 64) {
    const uint32_t t = s / per_tile, jj = s - t * per_tile;
    const uint64_t ta = tile0 + t;
    if (ta < ntiles) {
      const V v = *reinterpret_cast<const V*>(stage + t * LEN * 16 + jj * E);
      gptr<T> dst = gp(static_cast<T*>(out)) + ((ta * nq + q) * LEN) * 16 + jj * E;
      __builtin_nontemporal_store(v, reinterpret_cast<gptr<V>>(dst));
    }
  }
}

// Lane per cell over tiles of 16, whole wavefronts (four tiles): X, J and det J of each point
// go through LDS and leave as 16-byte stores, one point at a time. Lanes at or past the padded
// cell count compute the last padded lane's cell and write nothing of their own.
template <typename T, int TD, int GD, int K, bool SHIFT, typename TN = T>
__device__ __forceinline__ void direct_cell_staged(const TN* __restrict__ nodes,
                                                   const uint32_t* __restrict__ dofmap,
                                                   const uint32_t* __restrict__ flags,
                                                   const T* __restrict__ tab,
                                                   const double* __restrict__ tab64,
                                                   const T* __restrict__ ref_normals,
                                                   const FactorParams<T>& p,
                                                   const FactorOutputs& o, uint64_t cl_in,
                                                   T* __restrict__ stage) {
  constexpr int JL = TD * GD;
  const bool live = cl_in < p.padded_cells;
  const uint64_t cl = live ? cl_in : p.padded_cells - 1;
  const uint64_t c = cl < p.num_cells ? cl : p.num_cells - 1;
  const uint32_t* __restrict__ row = dofmap + c * K;
  T x[K][GD];
  gather_cell<T, GD, K, SHIFT, TN>(nodes, row, p.last_node, x);
  const uint32_t w = threadIdx.x & 63u, wt = w >> 4, wl = w & 15u;
  const uint64_t tile0 = (cl_in - w) >> 4, tile = cl >> 4;
  const uint64_t ntiles = p.padded_cells >> 4;
  const uint32_t mask = p.mask;
  const bool need_j = (mask & ~kFPoints) != 0;
  const uint32_t ex = 0;
  const uint32_t ej = ex + ((mask & kFPoints) ? 64u * GD : 0u);
  const uint32_t ed = ej + ((mask & kFJac) ? 64u * JL : 0u);
  const uint32_t own = live ? mask & ~(kFPoints | kFJac | kFDetJ) : 0u;
  for (uint32_t q = 0; q < p.nq; ++q) {
    T xv[GD], j[JL];
    contract<T, TD, GD, K, SHIFT>(x, tab + static_cast<uint64_t>(q) * K * 4,
                                  (mask & kFPoints) != 0, need_j, xv, j);
    if (mask & kFPoints) {
#pragma unroll
      for (int g = 0; g < GD; ++g) stage[ex + (wt * GD + g) * 16 + wl] = xv[g] + p.origin[g];
    }
    if (need_j) {
      const T detj = (mask & kFDetJ) ? detj_checked<T, TD, GD, K>(j, p, nodes, row, tab64, q) : T(0);
      if (mask & kFJac) {
#pragma unroll
        for (int i = 0; i < JL; ++i) stage[ej + (wt * JL + i) * 16 + wl] = j[i];
      }
      if (mask & kFDetJ) stage[ed + wt * 16 + wl] = detj;
      if (own) store_from_j<T, TD, GD>(j, detj, ref_normals, flags, c, q, own, tile * p.nq + q, 4u, wl, o);
    }
    wave_sync();
    if (mask & kFPoints) copy_point<T, GD>(stage + ex, o.x, tile0, ntiles, p.nq, q, w);
    if (mask & kFJac) copy_point<T, JL>(stage + ej, o.j, tile0, ntiles, p.nq, q, w);
    if (mask & kFDetJ) copy_point<T, 1>(stage + ed, o.detj, tile0, ntiles, p.nq, q, w);
    wave_sync();
  }
}

// Copies rows q0..q0+nv-1 of one staged field of a tile: they are contiguous (nv * LEN * 16
// values) in the staging buffer and in the QuadField, and leave as 16-byte stores.
template <typename T, int LEN>
__device__ __forceinline__ void copy_rows(const T* __restrict__ stage, void* out, uint64_t tile,
                                          uint32_t nq, uint32_t q0, uint32_t nv, uint32_t w) {
  using V = typename Vec16<T>::type;
  constexpr uint32_t E = 16 / sizeof(T);
  const uint32_t n = nv * (LEN * 16 / E);
  gptr<T> dst = gp(static_cast<T*>(out)) + ((tile * nq + q0) * LEN) * 16;
  for (uint32_t s = w; s < n; s += 64) {
    const V v = *reinterpret_cast<const V*>(stage + s * E);
    __builtin_nontemporal_store(v, reinterpret_cast<gptr<V>>(dst + s * E));
  }
}

// Gather rounds of a tile in the persistent kernel: four lanes per cell, at most kHybridMaxNodes
// nodes.
constexpr uint32_t kHybridGatherRounds = (kHybridMaxNodes + 3) / 4;

// The node coordinates of the 16 cells of a tile into LDS, xs[(k * GD + g) * 16 + cell]; with
// SHIFT nodes 1.. as offsets from node 0, as gather_cell takes them (nodes stored wider than T:
// the offsets taken in the storage type and rounded once). Lane = cell + 16 x (node group), at
// most 4 ROUNDS nodes per cell; every round loads (past the last node, node k - 1 again) before
// anything is stored, so the loads overlap.
template <typename T, int GD, bool SHIFT, uint32_t ROUNDS, typename TN = T>
__device__ __forceinline__ void gather_tile(const TN* __restrict__ nodes,
                                            const uint32_t* __restrict__ row, uint32_t k,
                                            uint32_t last_node, uint32_t w, T* __restrict__ xs) {
  const uint32_t cl16 = w & 15u, k0 = w >> 4;
  uint32_t idx[ROUNDS];
#pragma unroll
  for (uint32_t i = 0; i < ROUNDS; ++i) {
    const uint32_t kk = k0 + 4 * i;
    const uint32_t n = gp(row)[kk < k ? kk : k - 1];  // unconditional: the loads go together
    idx[i] = n < last_node ? n : last_node;
  }
  TN x0[GD];
  if constexpr (SHIFT) {
    const uint32_t r0 = gp(row)[0];
    const uint32_t n0 = r0 < last_node ? r0 : last_node;
#pragma unroll
    for (int g = 0; g < GD; ++g) x0[g] = gp(nodes)[static_cast<uint64_t>(n0) * GD + g];
  }
  TN v[ROUNDS][GD];
#pragma unroll
  for (uint32_t i = 0; i < ROUNDS; ++i)
#pragma unroll
    for (int g = 0; g < GD; ++g) v[i][g] = gp(nodes)[static_cast<uint64_t>(idx[i]) * GD + g];
#pragma unroll
  for (uint32_t i = 0; i < ROUNDS; ++i) {
    const uint32_t kk = k0 + 4 * i;
    if (kk < k) {
#pragma unroll
      for (int g = 0; g < GD; ++g) {
        if constexpr (sizeof(TN) != sizeof(T)) {
          xs[(kk * GD + g) * 16 + cl16] = static_cast<T>(kk > 0 ? v[i][g] - x0[g] : v[i][g]);
        } else {
          T x = v[i][g];
          if constexpr (SHIFT) {
            if (kk > 0) x -= x0[g];
          }
          xs[(kk * GD + g) * 16 + cl16] = x;
        }
      }
    }
  }
}

// All requested factors of the 16 cells of one tile (tiles of 16 cells), by one wavefront: the
// lanes are 16 cells x 4 point groups, and each step covers points q0..q0+3, whose rows of the
// tile are contiguous in every field. X, J and det J go through LDS and leave as sequential
// 16-byte non-temporal stores, so each wavefront streams through its tile's output in order;
// other factors are stored from registers.
template <typename T, int TD, int GD, int K, bool SHIFT, typename TN = T>
__device__ __forceinline__ void direct_tile(const TN* __restrict__ nodes,
                                            const uint32_t* __restrict__ dofmap,
                                            const uint32_t* __restrict__ flags,
                                            const T* __restrict__ tab,
                                            const double* __restrict__ tab64,
                                            const T* __restrict__ ref_normals,
                                            const FactorParams<T>& p, const FactorOutputs& o,
                                            uint64_t tile, T* __restrict__ stage,
                                            T* __restrict__ xs = nullptr) {
  constexpr int JL = TD * GD;
  const uint32_t w = threadIdx.x & 63u, cl16 = w & 15u, qg = w >> 4;
  const uint64_t cl = tile * 16 + cl16;
  const uint64_t c = cl < p.num_cells ? cl : p.num_cells - 1;
  const uint32_t* __restrict__ row = dofmap + c * K;
  T x[K][GD];
  if constexpr (sizeof(TN) != sizeof(T)) {
    // Wide node storage: the four lanes of a cell share its gather through LDS (xs, K x GD x 16
    // values) instead of each loading every node, which doubled the load traffic of FP64 nodes.
    gather_tile<T, GD, SHIFT, (K + 3) / 4, TN>(nodes, row, K, p.last_node, w, xs);
    wave_sync();
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int g = 0; g < GD; ++g) x[k][g] = xs[(k * GD + g) * 16 + cl16];
  } else {
    gather_cell<T, GD, K, SHIFT, TN>(nodes, row, p.last_node, x);
  }
  const uint32_t mask = p.mask;
  const bool need_j = (mask & ~kFPoints) != 0;
  const uint32_t ex = 0;
  const uint32_t ej = ex + ((mask & kFPoints) ? 64u * GD : 0u);
  const uint32_t ed = ej + ((mask & kFJac) ? 64u * JL : 0u);
  const uint32_t own = mask & ~(kFPoints | kFJac | kFDetJ);
  for (uint32_t q0 = 0; q0 < p.nq; q0 += 4) {
    const uint32_t nv = p.nq - q0 < 4u ? p.nq - q0 : 4u;
    const bool qlive = qg < nv;
    const uint32_t q = qlive ? q0 + qg : p.nq - 1;
    T xv[GD], j[JL];
    contract<T, TD, GD, K, SHIFT>(x, tab + static_cast<uint64_t>(q) * K * 4,
                                  (mask & kFPoints) != 0, need_j, xv, j);
    if (mask & kFPoints) {
#pragma unroll
      for (int g = 0; g < GD; ++g) stage[ex + (qg * GD + g) * 16 + cl16] = xv[g] + p.origin[g];
    }
    if (need_j) {
      const T detj = (mask & kFDetJ) ? detj_checked<T, TD, GD, K>(j, p, nodes, row, tab64, q) : T(0);
      if (mask & kFJac) {
#pragma unroll
        for (int i = 0; i < JL; ++i) stage[ej + (qg * JL + i) * 16 + cl16] = j[i];
      }
      if (mask & kFDetJ) stage[ed + qg * 16 + cl16] = detj;
      if (own && qlive)
        store_from_j<T, TD, GD>(j, detj, ref_normals, flags, c, q, own, tile * p.nq + q, 4u, cl16, o);
    }
    wave_sync();
    if (mask & kFPoints) copy_rows<T, GD>(stage + ex, o.x, tile, p.nq, q0, nv, w);
    if (mask & kFJac) copy_rows<T, JL>(stage + ej, o.j, tile, p.nq, q0, nv, w);
    if (mask & kFDetJ) copy_rows<T, 1>(stage + ed, o.detj, tile, p.nq, q0, nv, w);
    wave_sync();
  }
}

// Elements of the staging buffer per wavefront (X, J, det J of 64 (cell, point) pairs in 3D).
constexpr int kStageElems = 64 * 13;

// Copies n table values (a multiple of 4) from global memory into LDS with 16-byte accesses;
// every thread of the workgroup calls.
template <typename T>
__device__ __forceinline__ void table_to_lds(const T* __restrict__ src, T* __restrict__ dst,
                                             uint32_t n) {
  using V = typename Vec16<T>::type;
  constexpr uint32_t E = 16 / sizeof(T);
  for (uint32_t i = threadIdx.x; i < n / E; i += blockDim.x)
    reinterpret_cast<V*>(dst)[i] = reinterpret_cast<gptr<const V>>(gp(src))[i];
}

// One block. kMapTile: one wavefront per tile of 16 cells, the basis table staged in LDS;
// kMapCellStaged / kMapCell: one lane per cell, the table read through the scalar cache.
// Dynamic LDS: the staging buffers (staged mappings), then the table (kMapTile).
template <typename T, int TD, int GD, int K, bool SHIFT, typename TN = T>
__global__ void __launch_bounds__(kBlock)
    factor_direct(const TN* __restrict__ nodes, const uint32_t* __restrict__ dofmap,
                  const uint32_t* __restrict__ flags, const T* __restrict__ tab,
                  const double* __restrict__ tab64, const T* __restrict__ ref_normals,
                  FactorParams<T> p, FactorOutputs o) {
  extern __shared__ __align__(16) unsigned char tg_lds[];
  T* stage = reinterpret_cast<T*>(tg_lds);
  const uint32_t b = p.xcd_remap ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
  if (p.mapping == kMapTile) {
    T* ltab = stage + (kBlock / 64) * kStageElems;
    table_to_lds<T>(tab, ltab, p.nq * K * 4);
    __syncthreads();
    const uint64_t tile = static_cast<uint64_t>(b) * (kBlock / 64) + (threadIdx.x >> 6);
    if (tile >= (p.padded_cells >> 4)) return;
    // Wide node storage: each wavefront's gather buffer follows the table.
    T* xs = ltab + p.nq * K * 4 + (threadIdx.x >> 6) * (K * GD * 16);
    direct_tile<T, TD, GD, K, SHIFT, TN>(nodes, dofmap, flags, ltab, tab64, ref_normals, p, o, tile,
                                     stage + (threadIdx.x >> 6) * kStageElems, xs);
  } else if (p.mapping == kMapCellStaged) {
    const uint64_t cl = static_cast<uint64_t>(b) * kBlock + threadIdx.x;
    if (cl - (threadIdx.x & 63u) >= p.padded_cells) return;  // whole wavefronts stay
    direct_cell_staged<T, TD, GD, K, SHIFT, TN>(nodes, dofmap, flags, tab, tab64, ref_normals, p, o,
                                            cl, stage + (threadIdx.x >> 6) * kStageElems);
  } else {
    const uint64_t cl = static_cast<uint64_t>(b) * kBlock + threadIdx.x;
    if (cl >= p.padded_cells) return;
    direct_cell<T, TD, GD, K, SHIFT, TN>(nodes, dofmap, flags, tab, tab64, ref_normals, p, o, cl);
  }
}

// The shapes of the persistent mixed-mesh kernel, in the order of their shape index: every
// (tdim, gdim, k) with a register-resident kernel and k <= 27.
#define TG_HYBRID_SHAPES(X)                                                                    \
  X(0, 3, 3, 4) X(1, 3, 3, 10) X(2, 3, 3, 20) X(3, 3, 3, 8) X(4, 3, 3, 27) X(5, 3, 3, 6)     \
  X(6, 3, 3, 18) X(7, 3, 3, 5) X(8, 3, 3, 14) X(9, 2, 3, 3) X(10, 2, 3, 6) X(11, 2, 3, 10)    \
  X(12, 2, 3, 4) X(13, 2, 3, 9) X(14, 2, 2, 3) X(15, 2, 2, 6) X(16, 2, 2, 10) X(17, 2, 2, 4) \
  X(18, 2, 2, 9)

// All requested factors of one tile of 16 cells of a block, by one wavefront, with the node count
// k known only at run time: the cells' coordinates are gathered into LDS (gather_tile) and every
// sum reads them there, against the block's table in LDS. Lanes = 16 cells x 4 points, and X, J
// and det J leave through the staging buffer as direct_tile's do. The sums are those of
// contract_fixed term for term (first product, then fused multiply-adds in ascending node order),
// so the values equal the register-resident kernels' bit for bit; only the registers differ (a
// few dozen, whatever k).
template <typename T, int TD, int GD, int KC, bool SHIFT>
__device__ __forceinline__ void hybrid_tile(const HybridBlock<T>* __restrict__ hb,
                                            const T* __restrict__ ltab, uint64_t tile,
                                            T* __restrict__ xs, T* __restrict__ stage) {
  constexpr int JL = TD * GD;
  constexpr uint32_t K0 = SHIFT ? 1u : 0u;
  // The block's scalars and pointers, read once: the kernel's stores could otherwise make the
  // compiler read them again after every store.
  const FactorParams<T> p = hb->p;
  const FactorOutputs o = hb->out;
  const T* __restrict__ nodes = hb->nodes;
  const double* __restrict__ tab64 = hb->tab64;
  const T* __restrict__ ref_normals = hb->ref_normals;
  const uint32_t* __restrict__ flags = hb->flags;
  const uint32_t w = threadIdx.x & 63u, cl16 = w & 15u, qg = w >> 4;
  const uint32_t K = KC > 0 ? static_cast<uint32_t>(KC) : p.k;  // KC: the node count, if fixed
  const uint64_t cl = tile * 16 + cl16;
  const uint64_t c = cl < p.num_cells ? cl : p.num_cells - 1;
  const uint32_t* __restrict__ row = hb->dofmap + c * K;
  gather_tile<T, GD, SHIFT, kHybridGatherRounds>(nodes, row, K, p.last_node, w, xs);
  wave_sync();
  T x0[GD];
#pragma unroll
  for (int g = 0; g < GD; ++g) x0[g] = xs[g * 16 + cl16];
  const uint32_t mask = p.mask;
  const uint32_t ex = 0;
  const uint32_t ej = ex + ((mask & kFPoints) ? 64u * GD : 0u);
  const uint32_t ed = ej + ((mask & kFJac) ? 64u * JL : 0u);
  const uint32_t own = mask & ~(kFPoints | kFJac | kFDetJ);
  const bool need_j = (mask & ~kFPoints) != 0;
  for (uint32_t q0 = 0; q0 < p.nq; q0 += 4) {
    const uint32_t nv = p.nq - q0 < 4u ? p.nq - q0 : 4u;
    const bool qlive = qg < nv;
    const uint32_t q = qlive ? q0 + qg : p.nq - 1;
    const T* __restrict__ t = ltab + static_cast<uint64_t>(q) * K * 4;
    T ax[GD], aj[JL];
    {
      const TabRow<T> r = load_row<T>(t + 4 * K0);
#pragma unroll
      for (int g = 0; g < GD; ++g) {
        const T xk = xs[(K0 * GD + g) * 16 + cl16];
        ax[g] = xk * r.v[0];
#pragma unroll
        for (int d = 0; d < TD; ++d) aj[d * GD + g] = xk * r.v[1 + d];
      }
    }
#pragma unroll 3
    for (uint32_t k = K0 + 1; k < K; ++k) {
      const TabRow<T> r = load_row<T>(t + 4 * k);
#pragma unroll
      for (int g = 0; g < GD; ++g) {
        const T xk = xs[(k * GD + g) * 16 + cl16];
        ax[g] = tfma(xk, r.v[0], ax[g]);
#pragma unroll
        for (int d = 0; d < TD; ++d) aj[d * GD + g] = tfma(xk, r.v[1 + d], aj[d * GD + g]);
      }
    }
    if (mask & kFPoints) {
#pragma unroll
      for (int g = 0; g < GD; ++g) {
        const T xv = SHIFT ? x0[g] + ax[g] : ax[g];
        stage[ex + (qg * GD + g) * 16 + cl16] = xv + p.origin[g];
      }
    }
    if (need_j) {
      const T detj =
          (mask & kFDetJ) ? detj_checked_k<T, TD, GD>(aj, p, nodes, row, K, tab64, q) : T(0);
      if (mask & kFJac) {
#pragma unroll
        for (int i = 0; i < JL; ++i) stage[ej + (qg * JL + i) * 16 + cl16] = aj[i];
      }
      if (mask & kFDetJ) stage[ed + qg * 16 + cl16] = detj;
      if (own && qlive)
        store_from_j<T, TD, GD>(aj, detj, ref_normals, flags, c, q, own, tile * p.nq + q, 4u,
                                cl16, o);
    }
    wave_sync();
    if (mask & kFPoints) copy_rows<T, GD>(stage + ex, o.x, tile, p.nq, q0, nv, w);
    if (mask & kFJac) copy_rows<T, JL>(stage + ej, o.j, tile, p.nq, q0, nv, w);
    if (mask & kFDetJ) copy_rows<T, 1>(stage + ed, o.detj, tile, p.nq, q0, nv, w);
    wave_sync();
  }
}

// One persistent launch over the work items of every block of a mesh. An item is kHybridWaves
// tiles of 16 cells of one block, one per wavefront (hybrid_tile); the workgroup's wavefronts
// share the table in LDS. Items, in block order, are split on the
// host into one contiguous range of about equal work per workgroup, and the ranges of the
// workgroups of one XCD are adjacent (workgroup b runs on XCD b % 8), so each XCD walks one
// stretch of the mesh. No counters: the launch is deterministic and can be replayed from a graph
// as is. Dynamic LDS: the wavefronts' staging buffers, their coordinate buffers (xs_elems each)
// and the table of the block in hand.
template <typename T, bool SHIFT>
__global__ void __launch_bounds__(kHybridWaves * 64)
    factor_hybrid(const HybridBlock<T>* __restrict__ blocks, const WorkItem* __restrict__ items,
                  const uint32_t* __restrict__ ranges, uint32_t xs_elems) {
  extern __shared__ __align__(16) unsigned char tg_lds[];
  T* lds = reinterpret_cast<T*>(tg_lds);
  const uint32_t wv = threadIdx.x >> 6;
  T* stage = lds + wv * kStageElems;
  T* xs = lds + kHybridWaves * kStageElems + wv * xs_elems;
  T* ltab = lds + kHybridWaves * (kStageElems + xs_elems);
  uint32_t table_block = 0xffffffffu;
  const uint32_t r = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t begin = ranges[r], end = ranges[r + 1];
  for (uint32_t it = begin; it < end; ++it) {
    const WorkItem w = items[it];
    const HybridBlock<T>* __restrict__ hb = blocks + w.block;
    if (w.block != table_block) {
      // A new block's table: the workgroup is in step (items are uniform), so it reloads it.
      __syncthreads();
      table_to_lds<T>(hb->tab, ltab, hb->p.nq * hb->p.k * 4);
      __syncthreads();
      table_block = w.block;
    }
    const uint64_t tile = (static_cast<uint64_t>(w.first_cell) >> 4) + wv;
    if (tile >= (hb->p.padded_cells >> 4)) continue;
    // One body per (tdim, gdim), the node count read at run time: unrolling each shape's node
    // loop measured slower (more registers, longer code) than the runtime loop.
    uint32_t cls = 0;
    switch (hb->shape) {
#define TG_HCLASS(S_, TD_, GD_, K_) \
  case S_:                          \
    cls = TD_ * 4 + GD_;            \
    break;
      TG_HYBRID_SHAPES(TG_HCLASS)
#undef TG_HCLASS
      default:
        break;
    }
    if (cls == 3 * 4 + 3)
      hybrid_tile<T, 3, 3, 0, SHIFT>(hb, ltab, tile, xs, stage);
    else if (cls == 2 * 4 + 3)
      hybrid_tile<T, 2, 3, 0, SHIFT>(hb, ltab, tile, xs, stage);
    else if (cls == 2 * 4 + 2)
      hybrid_tile<T, 2, 2, 0, SHIFT>(hb, ltab, tile, xs, stage);
  }
}

// Runtime node count (the reference's `loops` strategy, layout_loops.rs:68-108): node
// coordinates are re-read through the dofmap at every point. Any element.
template <typename T, int TD, int GD, bool SHIFT, typename TN = T>
__global__ void __launch_bounds__(kBlock)
    factor_generic(const TN* __restrict__ nodes, const uint32_t* __restrict__ dofmap,
                   const uint32_t* __restrict__ flags, const T* __restrict__ tab,
                   const double* __restrict__ tab64, const T* __restrict__ ref_normals,
                   FactorParams<T> p, FactorOutputs o) {
  const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
  const uint64_t cl = static_cast<uint64_t>(b) * kBlock + threadIdx.x;
  if (cl >= p.padded_cells) return;
  const uint64_t c = cl < p.num_cells ? cl : p.num_cells - 1;
  const uint32_t K = p.k;
  const uint32_t* __restrict__ row = dofmap + c * K;
  TN x0[GD];  // node 0 in the storage type: FP64 storage takes the offsets in FP64
  {
    const uint32_t n = row[0] < p.last_node ? row[0] : p.last_node;
#pragma unroll
    for (int g = 0; g < GD; ++g) x0[g] = nodes[static_cast<uint64_t>(n) * GD + g];
  }
  const uint32_t k0 = SHIFT ? 1u : 0u;
  const uint32_t tl = p.tile_log2;
  const uint64_t tile = cl >> tl;
  const uint32_t lane = static_cast<uint32_t>(cl) & ((1u << tl) - 1u);
  const uint32_t mask = p.mask;
  const bool need_j = (mask & ~kFPoints) != 0;

  for (uint32_t q = 0; q < p.nq; ++q) {
    const T* __restrict__ tq = tab + static_cast<uint64_t>(q) * K * 4;
    T xs[GD];
    T j[TD * GD];
    for (uint32_t k = k0; k < K; ++k) {
      const uint32_t n = row[k] < p.last_node ? row[k] : p.last_node;
      T xk[GD];
#pragma unroll
      for (int g = 0; g < GD; ++g) {
        const TN v = nodes[static_cast<uint64_t>(n) * GD + g];
        xk[g] = static_cast<T>(SHIFT ? v - x0[g] : v);
      }
      const T* t = tq + 4 * k;
      if (k == k0) {
#pragma unroll
        for (int g = 0; g < GD; ++g) {
          xs[g] = xk[g] * t[0];
#pragma unroll
          for (int d = 0; d < TD; ++d) j[d * GD + g] = xk[g] * t[1 + d];
        }
      } else {
#pragma unroll
        for (int g = 0; g < GD; ++g) {
          xs[g] = tfma(xk[g], t[0], xs[g]);
#pragma unroll
          for (int d = 0; d < TD; ++d) j[d * GD + g] = tfma(xk[g], t[1 + d], j[d * GD + g]);
        }
      }
    }
    const uint64_t row_q = tile * p.nq + q;
    if (mask & kFPoints) {
#pragma unroll
      for (int g = 0; g < GD; ++g) {
        const T v = SHIFT ? (static_cast<T>(x0[g]) + xs[g]) + p.origin[g] : xs[g] + p.origin[g];
        st<T>(o.x, (((row_q * GD) + g) << tl) | lane, v);
      }
    }
    if (need_j) {
      T detj = T(0);
      if (mask & kFDetJ) {
        detj = det_of<T, TD, GD>(j);
        if constexpr (TD == GD && sizeof(T) == 4) {
          if (p.fixup) {
            T n2 = T(0);
#pragma unroll
            for (int i = 0; i < TD * GD; ++i) n2 = tfma(j[i], j[i], n2);
            if (__builtin_fabsf(detj) < static_cast<T>(p.tau) * n2 * tsqrt(n2))
              detj = static_cast<T>(detj_fp64<TD, GD>(nodes, row, K, p.last_node, tab64, q));
          }
        }
      }
      store_from_j<T, TD, GD>(j, detj, ref_normals, flags, c, q, mask, row_q, tl, lane, o);
    }
  }
}

// E-vector gather (evector.rs:197-236): out[(k * gdim + g) * padded + cell].
template <typename T>
__global__ void __launch_bounds__(kBlock)
    evector_kernel(const T* __restrict__ nodes, const uint32_t* __restrict__ dofmap,
                   uint64_t num_cells, uint64_t padded, uint32_t k, int gdim, uint32_t last_node,
                   T* __restrict__ out) {
  const uint64_t cl = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (cl >= padded) return;
  const uint64_t c = cl < num_cells ? cl : num_cells - 1;
  for (uint32_t kk = 0; kk < k; ++kk) {
    uint32_t n = dofmap[c * k + kk];
    n = n < last_node ? n : last_node;
    for (int g = 0; g < gdim; ++g)
      out[(static_cast<uint64_t>(kk) * gdim + g) * padded + cl] =
          nodes[static_cast<uint64_t>(n) * gdim + g];
  }
}

template <typename T, typename TN = T>
using FactorFn = void (*)(const TN*, const uint32_t*, const uint32_t*, const T*, const double*,
                          const T*, FactorParams<T>, FactorOutputs);

// Register-resident kernels exist for these (tdim, gdim, k): degrees 1-3 of every cell (tets,
// hexes up to Q3, prisms and pyramids up to degree 2, surfaces and curves up to degree 3).
#define TG_DIRECT_SHAPES(X)                                                                     \
  X(3, 3, 4) X(3, 3, 10) X(3, 3, 20) X(3, 3, 8) X(3, 3, 27) X(3, 3, 64) X(3, 3, 6) X(3, 3, 18) \
  X(3, 3, 5) X(3, 3, 14) X(2, 2, 3) X(2, 2, 6) X(2, 2, 10) X(2, 2, 4) X(2, 2, 9) X(2, 2, 16)   \
  X(2, 3, 3) X(2, 3, 6) X(2, 3, 10) X(2, 3, 4) X(2, 3, 9) X(2, 3, 16) X(1, 1, 2) X(1, 1, 3)    \
  X(1, 1, 4) X(1, 2, 2) X(1, 2, 3) X(1, 2, 4) X(1, 3, 2) X(1, 3, 3) X(1, 3, 4)

template <typename T, typename TN>
FactorFn<T, TN> direct_kernel(int td, int gd, uint32_t k, bool shift) {
  // FP64 keeps at most 27 nodes in registers; larger elements use the runtime-K kernel.
  if (sizeof(T) == 8 && k > 27) return nullptr;
  if (sizeof(T) == 4 && !shift) return nullptr;  // F32 without the shift runs generic
  if constexpr (sizeof(TN) != sizeof(T)) {  // wider node storage: only with the shift
#define TG_CASE(TD_, GD_, K_) \
  if (td == TD_ && gd == GD_ && k == K_) return factor_direct<T, TD_, GD_, K_, true, TN>;
    TG_DIRECT_SHAPES(TG_CASE)
#undef TG_CASE
  } else {
#define TG_CASE(TD_, GD_, K_)                                                       \
  if (td == TD_ && gd == GD_ && k == K_)                                            \
    return shift ? static_cast<FactorFn<T, TN>>(factor_direct<T, TD_, GD_, K_, true>) \
                 : static_cast<FactorFn<T, TN>>(factor_direct<T, TD_, GD_, K_, false>);
    TG_DIRECT_SHAPES(TG_CASE)
#undef TG_CASE
  }
  return nullptr;
}

// The `cooperative` strategy (tetraloma-geometry/src/cooperative/mod.rs:1-33, factors.rs:1-183):
// each sum over a cell's k nodes is split among L lanes, L the largest power of two not above
// min(32, k), so no lane is empty. Leaf l sums the nodes l, l + L, l + 2L, ... in ascending
// order from the product of node l; the leaves are then added by the reference's pairwise tree
// (at strides t = 1, 2, ..., leaf u with u mod 2t = 0 adds leaf u + t's partial, left + right).
// A cell's L lanes are consecutive lanes of one wavefront, so the tree runs on cross-lane
// shuffles, and the cell's first lane forms and stores every requested factor. FP32 works on
// offsets from the cell's node 0 (node 0 itself then adds zero), X adds node 0 back. For
// high-degree cells, whose nodes do not fit one lane's registers and which one lane per cell
// would sum in a long series.
constexpr uint32_t kCoopRounds = 8;  // nodes per leaf: at most 8 x 32 = 256 nodes per cell

template <typename T, int TD, int GD, bool SHIFT>
__global__ void __launch_bounds__(kBlock)
    factor_coop(const T* __restrict__ nodes, const uint32_t* __restrict__ dofmap,
                const uint32_t* __restrict__ flags, const T* __restrict__ tab,
                const double* __restrict__ tab64, const T* __restrict__ ref_normals,
                FactorParams<T> p, FactorOutputs o, uint32_t leaves) {
  constexpr int JL = TD * GD;
  const uint32_t L = leaves;  // a power of two, 1..32
  const uint32_t leaf = threadIdx.x & (L - 1);
  const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
  const uint64_t cl = static_cast<uint64_t>(b) * (kBlock / L) + threadIdx.x / L;
  if (cl >= p.padded_cells) return;  // a cell's L lanes leave together
  const uint64_t c = cl < p.num_cells ? cl : p.num_cells - 1;
  const uint32_t K = p.k;
  const uint32_t* __restrict__ row = dofmap + c * K;
  T x0[GD];
  {
    const uint32_t n0 = row[0] < p.last_node ? row[0] : p.last_node;
#pragma unroll
    for (int g = 0; g < GD; ++g) x0[g] = nodes[static_cast<uint64_t>(n0) * GD + g];
  }
  // The leaf's nodes (past the last node, node k - 1 again: loaded, never summed).
  T x[kCoopRounds][GD];
#pragma unroll
  for (uint32_t i = 0; i < kCoopRounds; ++i) {
    const uint32_t k = leaf + i * L;
    const uint32_t r = row[k < K ? k : K - 1];
    const uint32_t n = r < p.last_node ? r : p.last_node;
#pragma unroll
    for (int g = 0; g < GD; ++g) {
      const T v = nodes[static_cast<uint64_t>(n) * GD + g];
      x[i][g] = SHIFT ? v - x0[g] : v;
    }
  }
  const uint32_t tl = p.tile_log2;
  const uint64_t tile = cl >> tl;
  const uint32_t lane = static_cast<uint32_t>(cl) & ((1u << tl) - 1u);
  const uint32_t mask = p.mask;
  const bool need_j = (mask & ~kFPoints) != 0;
  for (uint32_t q = 0; q < p.nq; ++q) {
    const T* __restrict__ t = tab + static_cast<uint64_t>(q) * K * 4;
    T ax[GD], aj[JL];
    {
      const TabRow<T> r = load_row<T>(t + 4 * leaf);
#pragma unroll
      for (int g = 0; g < GD; ++g) {
        ax[g] = x[0][g] * r.v[0];
#pragma unroll
        for (int d = 0; d < TD; ++d) aj[d * GD + g] = x[0][g] * r.v[1 + d];
      }
    }
#pragma unroll
    for (uint32_t i = 1; i < kCoopRounds; ++i) {
      const uint32_t k = leaf + i * L;
      if (k < K) {
        const TabRow<T> r = load_row<T>(t + 4 * k);
#pragma unroll
        for (int g = 0; g < GD; ++g) {
          ax[g] = tfma(x[i][g], r.v[0], ax[g]);
#pragma unroll
          for (int d = 0; d < TD; ++d) aj[d * GD + g] = tfma(x[i][g], r.v[1 + d], aj[d * GD + g]);
        }
      }
    }
    for (uint32_t st = 1; st < L; st <<= 1) {
      const bool left = (leaf & (2 * st - 1)) == 0;
#pragma unroll
      for (int g = 0; g < GD; ++g) {
        const T right = __shfl_down(ax[g], st, L);
        if (left) ax[g] = ax[g] + right;
      }
#pragma unroll
      for (int i = 0; i < JL; ++i) {
        const T right = __shfl_down(aj[i], st, L);
        if (left) aj[i] = aj[i] + right;
      }
    }
    if (leaf == 0) {
      const uint64_t row_q = tile * p.nq + q;
      if (mask & kFPoints) {
#pragma unroll
        for (int g = 0; g < GD; ++g) {
          const T xv = SHIFT ? x0[g] + ax[g] : ax[g];
          st<T>(o.x, (((row_q * GD) + g) << tl) | lane, xv + p.origin[g]);
        }
      }
      if (need_j) {
        const T detj =
            (mask & kFDetJ) ? detj_checked_k<T, TD, GD>(aj, p, nodes, row, K, tab64, q) : T(0);
        store_from_j<T, TD, GD>(aj, detj, ref_normals, flags, c, q, mask, row_q, tl, lane, o);
      }
    }
  }
}

// A product rounded before it is added to the running sum, as the measure's cell sums read
// (measure.rs:17-41): the first point starts the sum.
template <typename T>
__device__ __forceinline__ T weighted_sum(T acc, T d, T w, bool first) {
#pragma clang fp contract(off)
  const T t = d * w;
  return first ? t : acc + t;
}

// The measure's first blocked sums straight from each point's det J (tg_measure_block): one lane
// per cell computes det J at its points as the factor kernels do (detj_checked after contract)
// and sums them in point order with the weights narrowed to T (measure.rs:17-41); each 64 cells
// of the workgroup are one block of the first pass of 64-value sums (measure.rs:46-72), which
// one lane adds in cell order from LDS into partials[c / 64]. No det J field is written.
template <typename T, int TD, int GD, int K, bool SHIFT>
__global__ void __launch_bounds__(kBlock)
    measure_cells(const T* __restrict__ nodes, const uint32_t* __restrict__ dofmap,
                  const T* __restrict__ tab, const double* __restrict__ tab64,
                  const double* __restrict__ weights, FactorParams<T> p, T* __restrict__ partials) {
  // Each XCD takes one contiguous range of cells, so nodes shared by neighbouring cells are
  // fetched into one L2.
  constexpr uint32_t kRow = 65;  // LDS row of one block of 64 sums, padded against conflicts
  __shared__ T cell_sum[(kBlock / 64) * kRow];
  const uint32_t b = p.xcd_remap ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
  const uint64_t c = static_cast<uint64_t>(b) * kBlock + threadIdx.x;
  T acc = T(0);
  if (c < p.num_cells) {
    const uint32_t* __restrict__ row = dofmap + c * K;
    T x[K][GD];
    gather_cell<T, GD, K, SHIFT>(nodes, row, p.last_node, x);
    for (uint32_t q = 0; q < p.nq; ++q) {
      T xv[GD], j[TD * GD];
      contract<T, TD, GD, K, SHIFT>(x, tab + static_cast<uint64_t>(q) * K * 4, false, true, xv, j);
      const T d = detj_checked<T, TD, GD, K>(j, p, nodes, row, tab64, q);
      acc = weighted_sum<T>(acc, d, static_cast<T>(weights[q]), q == 0);
    }
  }
  cell_sum[(threadIdx.x >> 6) * kRow + (threadIdx.x & 63u)] = acc;
  __syncthreads();
  if (threadIdx.x < kBlock / 64) {
    const uint64_t first = static_cast<uint64_t>(b) * kBlock + threadIdx.x * 64u;
    if (first < p.num_cells) {
      const uint64_t left = p.num_cells - first;
      const uint32_t count = left < 64 ? static_cast<uint32_t>(left) : 64u;
      const T* __restrict__ row = cell_sum + threadIdx.x * kRow;
      T sum = row[0];
      for (uint32_t l = 1; l < count; ++l) sum += row[l];
      partials[first >> 6] = sum;
    }
  }
}

template <typename T>
using MeasureFn = void (*)(const T*, const uint32_t*, const T*, const double*, const double*,
                           FactorParams<T>, T*);

template <typename T>
MeasureFn<T> measure_kernel(int td, int gd, uint32_t k, bool shift) {
  if (sizeof(T) == 8 && k > 27) return nullptr;
  if (sizeof(T) == 4 && !shift) return nullptr;  // as direct_kernel: F32 without the shift is generic
#define TG_CASE(TD_, GD_, K_)                                                      \
  if (td == TD_ && gd == GD_ && k == K_)                                           \
    return shift ? static_cast<MeasureFn<T>>(measure_cells<T, TD_, GD_, K_, true>) \
                 : static_cast<MeasureFn<T>>(measure_cells<T, TD_, GD_, K_, false>);
  TG_DIRECT_SHAPES(TG_CASE)
#undef TG_CASE
  return nullptr;
}

template <typename T, typename TN>
FactorFn<T, TN> generic_kernel(int td, int gd, bool shift) {
  if constexpr (sizeof(TN) != sizeof(T)) {
    if (!shift) return nullptr;
#define TG_G(TD_, GD_) \
  if (td == TD_ && gd == GD_) return factor_generic<T, TD_, GD_, true, TN>;
    TG_G(1, 1) TG_G(1, 2) TG_G(1, 3) TG_G(2, 2) TG_G(2, 3) TG_G(3, 3)
#undef TG_G
  } else {
#define TG_G(TD_, GD_)                                                              \
  if (td == TD_ && gd == GD_)                                                       \
    return shift ? static_cast<FactorFn<T, TN>>(factor_generic<T, TD_, GD_, true>)  \
                 : static_cast<FactorFn<T, TN>>(factor_generic<T, TD_, GD_, false>);
    TG_G(1, 1) TG_G(1, 2) TG_G(1, 3) TG_G(2, 2) TG_G(2, 3) TG_G(3, 3)
#undef TG_G
  }
  return nullptr;
}

template <typename T, typename TN>
hipError_t launch_factor_impl(int tdim, int gdim, bool shift, bool generic, const TN* nodes,
                              const uint32_t* dofmap, const uint32_t* flags, const T* tab,
                              const double* tab64, const T* ref_normals,
                              const FactorParams<T>& pin, const FactorOutputs& out,
                              hipStream_t stream) {
  FactorFn<T, TN> fn = generic ? nullptr : direct_kernel<T, TN>(tdim, gdim, pin.k, shift);
  const bool direct = fn != nullptr;
  if (fn == nullptr) fn = generic_kernel<T, TN>(tdim, gdim, shift);
  if (fn == nullptr) return hipErrorNotSupported;
  FactorParams<T> p = pin;
  // The staged mappings need tiles of 16 and a register-resident kernel; kMapTile a table that
  // fits its LDS budget.
  if (!direct || p.tile_log2 != 4) p.mapping = kMapCell;
  const size_t table_bytes = static_cast<size_t>(p.nq) * p.k * 4 * sizeof(T);
  if (p.mapping == kMapTile && table_bytes > kMaxLdsTableBytes) p.mapping = kMapCellStaged;
  const size_t stage_bytes = (kBlock / 64) * kStageElems * sizeof(T);
  const size_t gather_bytes =
      sizeof(TN) != sizeof(T) ? (kBlock / 64) * static_cast<size_t>(p.k) * gdim * 16 * sizeof(T)
                              : 0;
  const size_t lds = p.mapping == kMapTile         ? stage_bytes + table_bytes + gather_bytes
                     : p.mapping == kMapCellStaged ? stage_bytes
                                                   : 0;
  const uint64_t per_block = cells_per_workgroup(p.mapping);
  const uint64_t blocks = (p.padded_cells + per_block - 1) / per_block;
  if (blocks > 0x7fffffffu) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(fn, dim3(static_cast<uint32_t>(blocks)), dim3(kBlock),
                     static_cast<uint32_t>(lds), stream, nodes, dofmap, flags, tab, tab64,
                     ref_normals, p, out);
  return hipGetLastError();
}

}  // namespace

#ifndef TG_FACTOR_NODES64
template <typename T>
hipError_t launch_factor_coop(int tdim, int gdim, bool shift, const T* nodes,
                              const uint32_t* dofmap, const uint32_t* flags, const T* tab,
                              const double* tab64, const T* ref_normals, const FactorParams<T>& p,
                              const FactorOutputs& out, hipStream_t stream) {
  using Fn = void (*)(const T*, const uint32_t*, const uint32_t*, const T*, const double*,
                      const T*, FactorParams<T>, FactorOutputs, uint32_t);
  Fn fn = nullptr;
#define TG_C(TD_, GD_)                                                                \
  if (tdim == TD_ && gdim == GD_)                                                     \
    fn = shift ? static_cast<Fn>(factor_coop<T, TD_, GD_, true>)                      \
               : static_cast<Fn>(factor_coop<T, TD_, GD_, false>);
  TG_C(1, 1) TG_C(1, 2) TG_C(1, 3) TG_C(2, 2) TG_C(2, 3) TG_C(3, 3)
#undef TG_C
  if (fn == nullptr || p.k == 0) return hipErrorNotSupported;
  uint32_t leaves = 1;
  while (leaves * 2 <= std::min<uint32_t>(32u, p.k)) leaves *= 2;
  if (p.k > leaves * kCoopRounds) return hipErrorNotSupported;
  const uint64_t per_block = kBlock / leaves;
  const uint64_t blocks = (p.padded_cells + per_block - 1) / per_block;
  if (blocks > 0x7fffffffu) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(fn, dim3(static_cast<uint32_t>(blocks)), dim3(kBlock), 0, stream, nodes,
                     dofmap, flags, tab, tab64, ref_normals, p, out, leaves);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_factor(int tdim, int gdim, bool shift, bool generic, const T* nodes,
                         const uint32_t* dofmap, const uint32_t* flags, const T* tab,
                         const double* tab64, const T* ref_normals, const FactorParams<T>& pin,
                         const FactorOutputs& out, hipStream_t stream) {
  return launch_factor_impl<T, T>(tdim, gdim, shift, generic, nodes, dofmap, flags, tab, tab64,
                                  ref_normals, pin, out, stream);
}
#else
hipError_t launch_factor_f64_nodes(int tdim, int gdim, bool generic, const double* nodes,
                                   const uint32_t* dofmap, const uint32_t* flags,
                                   const float* tab, const double* tab64,
                                   const float* ref_normals, const FactorParams<float>& p,
                                   const FactorOutputs& out, hipStream_t stream) {
  return launch_factor_impl<float, double>(tdim, gdim, true, generic, nodes, dofmap, flags, tab,
                                           tab64, ref_normals, p, out, stream);
}
#endif

template <typename T>
bool has_measure_cells(int tdim, int gdim, uint32_t k, bool shift) {
  return measure_kernel<T>(tdim, gdim, k, shift) != nullptr;
}

template <typename T>
bool has_direct_factor(int tdim, int gdim, uint32_t k, bool shift) {
  return direct_kernel<T, T>(tdim, gdim, k, shift) != nullptr;
}

template <typename T>
hipError_t launch_measure_cells(int tdim, int gdim, bool shift, const T* nodes,
                                const uint32_t* dofmap, const T* tab, const double* tab64,
                                const double* weights, const FactorParams<T>& p, T* partials,
                                hipStream_t stream) {
  MeasureFn<T> fn = measure_kernel<T>(tdim, gdim, p.k, shift);
  if (fn == nullptr) return hipErrorNotSupported;
  const uint64_t blocks = (p.num_cells + kBlock - 1) / kBlock;
  if (blocks == 0) return hipSuccess;
  if (blocks > 0x7fffffffu) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(fn, dim3(static_cast<uint32_t>(blocks)), dim3(kBlock), 0, stream, nodes,
                     dofmap, tab, tab64, weights, p, partials);
  return hipGetLastError();
}

#ifdef TG_FACTOR_COMMON  // type-independent helpers: defined in one of the two objects
uint32_t hybrid_xs_elems(uint32_t max_k, int gdim) {
  return (max_k * static_cast<uint32_t>(gdim) * 16 + 3) / 4 * 4;
}

size_t hybrid_lds_bytes(size_t elem, uint32_t xs_elems, size_t max_table_bytes) {
  return kHybridWaves * (kStageElems + xs_elems) * elem + max_table_bytes;
}

int hybrid_shape(int tdim, int gdim, uint32_t k) {
#define TG_HSHAPE(S_, TD_, GD_, K_) \
  if (tdim == TD_ && gdim == GD_ && k == K_) return S_;
  TG_HYBRID_SHAPES(TG_HSHAPE)
#undef TG_HSHAPE
  return -1;
}
#endif

template <typename T>
static auto hybrid_fn(bool shift) {
  return shift ? factor_hybrid<T, true> : factor_hybrid<T, false>;
}

template <typename T>
hipError_t hybrid_grid(bool shift, size_t lds, uint32_t* grid) {
  // Resident workgroups per CU from the kernel's registers and LDS: 512 VGPRs per SIMD lane,
  // kHybridWaves / 4 wavefronts of each workgroup per SIMD, 160 KiB of LDS per CU, at most 8
  // wavefronts per SIMD.
  int dev = 0, cus = 0;
  hipFuncAttributes attr{};
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess)
    e = hipFuncGetAttributes(&attr, reinterpret_cast<const void*>(hybrid_fn<T>(shift)));
  if (e != hipSuccess) return e;
  // The dynamic LDS must fit one workgroup; past the default limit the kernel opts in.
  int lds_default = 0, lds_optin = 0;
  e = hipDeviceGetAttribute(&lds_default, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&lds_optin, hipDeviceAttributeSharedMemPerBlockOptin, dev);
  if (e != hipSuccess) return e;
  const size_t lds_limit = static_cast<size_t>(std::max(lds_default, lds_optin));
  if (lds + attr.sharedSizeBytes > lds_limit) return hipErrorInvalidValue;
  if (lds + attr.sharedSizeBytes > static_cast<size_t>(lds_default)) {
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(hybrid_fn<T>(shift)),
                            hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return e;
  }
  const int regs = std::max(8, (attr.numRegs + 7) / 8 * 8);
  const int by_regs = std::min(8, 512 / regs) / (kHybridWaves / 4);
  const size_t per_wg = lds + attr.sharedSizeBytes;
  const int by_lds = per_wg ? static_cast<int>((160 * 1024) / per_wg) : 8;
  const int per_cu = std::max(1, std::min(by_regs, by_lds));
  *grid = static_cast<uint32_t>(std::max(1, cus) * per_cu);
  return hipSuccess;
}

template <typename T>
hipError_t launch_hybrid(bool shift, const HybridBlock<T>* blocks, const WorkItem* items,
                         const uint32_t* ranges, uint32_t grid, size_t lds, uint32_t xs_elems,
                         hipStream_t stream) {
  hipLaunchKernelGGL(hybrid_fn<T>(shift), dim3(grid), dim3(kHybridWaves * 64),
                     static_cast<uint32_t>(lds),
                     stream, blocks, items, ranges, xs_elems);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_evector(const T* nodes, const uint32_t* dofmap, uint64_t num_cells,
                          uint64_t padded, uint32_t k, int gdim, uint32_t last_node, T* out,
                          hipStream_t stream) {
  const uint64_t blocks = (padded + kBlock - 1) / kBlock;
  if (blocks > 0x7fffffffu) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(evector_kernel<T>, dim3(static_cast<uint32_t>(blocks)), dim3(kBlock), 0,
                     stream, nodes, dofmap, num_cells, padded, k, gdim, last_node, out);
  return hipGetLastError();
}

// The build compiles this file once per coordinate type (TG_FACTOR_TYPE = float or double) so
// that the two halves of the kernel set compile in parallel.
#ifndef TG_FACTOR_TYPE
#error "compile with -DTG_FACTOR_TYPE=float or -DTG_FACTOR_TYPE=double"
#endif
#ifndef TG_FACTOR_NODES64
using FT = TG_FACTOR_TYPE;
template hipError_t hybrid_grid<FT>(bool, size_t, uint32_t*);
template hipError_t launch_hybrid<FT>(bool, const HybridBlock<FT>*, const WorkItem*,
                                      const uint32_t*, uint32_t, size_t, uint32_t, hipStream_t);
template hipError_t launch_factor<FT>(int, int, bool, bool, const FT*, const uint32_t*,
                                      const uint32_t*, const FT*, const double*, const FT*,
                                      const FactorParams<FT>&, const FactorOutputs&, hipStream_t);
template hipError_t launch_factor_coop<FT>(int, int, bool, const FT*, const uint32_t*,
                                           const uint32_t*, const FT*, const double*, const FT*,
                                           const FactorParams<FT>&, const FactorOutputs&,
                                           hipStream_t);
template hipError_t launch_evector<FT>(const FT*, const uint32_t*, uint64_t, uint64_t, uint32_t,
                                       int, uint32_t, FT*, hipStream_t);
template bool has_measure_cells<FT>(int, int, uint32_t, bool);
template bool has_direct_factor<FT>(int, int, uint32_t, bool);
template hipError_t launch_measure_cells<FT>(int, int, bool, const FT*, const uint32_t*, const FT*,
                                             const double*, const double*,
                                             const FactorParams<FT>&, FT*, hipStream_t);
#endif

}  // namespace tg
