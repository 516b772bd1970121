// X, J, det J, K and facet normals of 3D cells of dense bases (20 nodes and more: P3 and higher
// tetrahedra, pyramids and prisms, and Q2 and higher hexahedra at point sets the sum-factorised
// kernels do not take), FP32 and FP64, with the node sums on the matrix cores
// (v_mfma_f32_16x16x4f32, v_mfma_f64_16x16x4f64). Same factors as the register-resident and
// runtime-K kernels
// (the reference's factor kernel, tetraloma-geometry/src/layout.rs:138-213: X_g = sum_k x_kg
// phi_k, J_gd = sum_k x_kg d_d phi_k, det J by first-row cofactors), with the sums taken as
// matrix products over the cells of a tile:
//
//   D_v[q][cell] = sum_k T_v[q][k] x_g[k][cell]      v = 0 (phi), 1..3 (d phi / d xi_d)
//
// One wavefront per tile of 16 cells. Its lanes gather the tile's nodes, four lanes per cell:
// lane l loads nodes l / 16, l / 16 + 4, ... of cell l % 16, which are its B operands (node
// 4s + l / 16 at step s), formed in registers as offsets from each cell's node 0 (zero for node
// 0 and past k; node 0 itself stays absolute and X adds it back, as every FP32 kernel here) and
// held for the whole tile; the A
// operand is the table value T_v[q = 16 qb + l % 16][k = 4s + l / 16], laid out on the host in
// operand order (one coalesced load per product from the L1-resident table). Each block of 16
// points takes 4 kinds x 3 coordinates x ceil(k / 4) products, and lane l receives the points
// 16 qb + 4 (l / 16) + i (i = 0..3, one per register) of cell l % 16: X and all of J of its
// points, so it forms det J (with the FP64 recomputation where |det J| < tau |J|_F^3). The
// block's rows of the tile are contiguous in each field: they are staged in LDS and leave as
// 16-byte non-temporal stores.
//
// Each step adds its four products to the accumulator in node order, each rounded once, as the
// register and runtime-K kernels' chain of fused multiply-adds from node 0 does: the values
// equal theirs bit for bit (measured on every shape and point set of the parity tests and on
// 663,552 P3 tetrahedra at 14 and 64 points), and all are held to the oracle. (The cooperative
// kernel adds its leaves by a tree and differs in the last bits.)
// Workgroups are persistent: those of one XCD walk one contiguous range of tiles, four tiles
// (one per wavefront) at a time. On FP64-stored nodes (TN = double) the offsets are taken in
// FP64 and rounded once to FP32, and the FP64 recomputation reads the FP64 nodes, as the
// register kernels on FP64 nodes do; the products are the same FP32 products.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dense_mfma.hpp"
#include "geom_internal.hpp"
#include "jacobian.hpp"
#include "sumfact_tile.hpp"

namespace tg {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;
// The factors formed from J: the J products run only for calls that ask for one of them.
constexpr uint32_t kNeedJ = kFJac | kFDetJ | kFPinv | kFFacetNormals;

__device__ __forceinline__ v4f mfma4(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// FP64: the same 16 x 16 x 4 shape and lane layout, one product of doubles per lane.
__device__ __forceinline__ v4d mfma4(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <typename T>
struct Acc;
template <>
struct Acc<float> {
  typedef v4f type;
};
template <>
struct Acc<double> {
  typedef v4d type;
};

template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> gp(T* p) {
  return (gptr<T>)p;
}

// acc[v][g] += sum over node steps s0 ... s0 + N - 1 (those below S) of the products of the A
// operands T_v (table row (v * S + s) * 64 from ta) with the B operands b[s - s0][g], in step
// order, for the kinds the call asks for (WX: v = 0, X; WJ: v = 1..3, J). The kinds are fixed at
// compile time so that the table loads carry no branches: they go in groups of two steps, each
// group's loads issued before the previous group's products, so a block waits about one round
// trip to the cache for its table instead of one a step.
template <bool WX, bool WJ, int N, typename T, typename V>
__device__ __forceinline__ void step_products(gptr<const T> ta, uint32_t S, uint32_t s0,
                                              const T (&b)[N][3], V (&acc)[4][3]) {
  constexpr int G = 2, NG = (N + G - 1) / G;
  T a[2][4][G];
  const auto load = [&](int gi) {
#pragma unroll
    for (int i = 0; i < G; ++i) {
      // Steps past S (and past N) load the last step's row: in bounds, never used.
      const uint32_t s0i = s0 + static_cast<uint32_t>(gi * G + i < N ? gi * G + i : N - 1);
      const uint32_t s = s0i < S ? s0i : S - 1;
#pragma unroll
      for (int vv = 0; vv < 4; ++vv)
        if (vv == 0 ? WX : WJ) a[gi & 1][vv][i] = ta[(vv * S + s) * 64];
    }
  };
  load(0);
#pragma unroll
  for (int gi = 0; gi < NG; ++gi) {
    if (gi + 1 < NG) load(gi + 1);
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const int si = gi * G + i;
      if (si >= N || s0 + static_cast<uint32_t>(si) >= S) break;  // uniform
#pragma unroll
      for (int vv = 0; vv < 4; ++vv)
        if (vv == 0 ? WX : WJ)
#pragma unroll
          for (int g = 0; g < 3; ++g) acc[vv][g] = mfma4(a[gi & 1][vv][i], b[si][g], acc[vv][g]);
    }
  }
}

// BATCH false: one table load per step and kind, each waited for before its products (the
// kinds tested at run time; fewer registers).
template <bool BATCH, int N, typename T, typename V>
__device__ __forceinline__ void products(uint32_t mask, gptr<const T> ta, uint32_t S, uint32_t s0,
                                         const T (&b)[N][3], V (&acc)[4][3]) {
  if constexpr (!BATCH) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if (s0 + i >= S) break;  // uniform
#pragma unroll
      for (int vv = 0; vv < 4; ++vv) {
        if (vv == 0 && !(mask & kFPoints)) continue;
        if (vv > 0 && !(mask & kNeedJ)) continue;
        const T a = ta[(vv * S + s0 + i) * 64];
#pragma unroll
        for (int g = 0; g < 3; ++g) acc[vv][g] = mfma4(a, b[i][g], acc[vv][g]);
      }
    }
    return;
  }
  const bool wx = (mask & kFPoints) != 0, wj = (mask & kNeedJ) != 0;
  if (wx && wj)
    step_products<true, true>(ta, S, s0, b, acc);
  else if (wx)
    step_products<true, false>(ta, S, s0, b, acc);
  else
    step_products<false, true>(ta, S, s0, b, acc);
}

// Elements of a wavefront's LDS: the stage of one block of 16 points, J [16][9][16], then, over
// it once J is stored, X [16][3][16] and det J [16][16] (36 KiB a workgroup: four per CU).
constexpr int kStageElems = 9 * 256;

// Copies n values (16 bytes of them a multiple) from LDS to global memory with 16-byte
// non-temporal stores spread over the wavefront's lanes.
template <typename T>
__device__ __forceinline__ void flush(const T* __restrict__ src, gptr<T> dst, uint32_t n,
                                      uint32_t lane) {
  constexpr uint32_t E = 16 / sizeof(T);
  typedef T v16 __attribute__((ext_vector_type(E)));
  for (uint32_t e = lane; e < n / E; e += 64u) {
    const v16 v = *reinterpret_cast<const v16*>(src + e * E);
    __builtin_nontemporal_store(v, reinterpret_cast<gptr<v16>>(dst + e * E));
  }
}

// FP64 det J of a cell of k nodes at point q from the stored coordinates (of type TN; offsets
// from node 0 in FP64), for near-degenerate points: the register and runtime-K kernels' sums, in
// node order.
template <typename TN>
__device__ __noinline__ double detj_fp64_k(const TN* __restrict__ nodes,
                                           const uint32_t* __restrict__ row, uint32_t k,
                                           uint32_t last_node, const double* __restrict__ tab64,
                                           uint32_t q) {
  double x0[3], j[9];
  const uint32_t n0 = row[0] < last_node ? row[0] : last_node;
#pragma unroll
  for (int g = 0; g < 3; ++g) x0[g] = static_cast<double>(nodes[static_cast<uint64_t>(n0) * 3 + g]);
#pragma unroll
  for (int i = 0; i < 9; ++i) j[i] = 0.0;
  for (uint32_t kk = 1; kk < k; ++kk) {
    const uint32_t n = row[kk] < last_node ? row[kk] : last_node;
    const double* t = tab64 + (static_cast<uint64_t>(q) * k + kk) * 4;
#pragma unroll
    for (int g = 0; g < 3; ++g) {
      const double d = static_cast<double>(nodes[static_cast<uint64_t>(n) * 3 + g]) - x0[g];
#pragma unroll
      for (int dd = 0; dd < 3; ++dd) j[dd * 3 + g] = __builtin_fma(d, t[1 + dd], j[dd * 3 + g]);
    }
  }
  return j[0] * (j[4] * j[8] - j[7] * j[5]) - j[3] * (j[1] * j[8] - j[7] * j[2]) +
         j[6] * (j[1] * j[5] - j[4] * j[2]);
}

// The end of one block of 16 points of a tile, from the products acc[v][g] (v = 0: X, as
// offsets from node 0 in FP32 and absolute in FP64; 1..3: J(g, d = v - 1)); lane (cell, kg) =
// (l % 16, l / 16) holds four rows qq of the block (points 16 qb + qq) of its cell: 4 kg + i in
// FP32, kg + 4 i in FP64 (the layouts of the two instructions' results). det J of every point
// first (in FP32 with the test for the FP64 recomputation); J, then K, go through the whole
// stage, the facet normals (from K and the point's reference normal) wait in registers; then X,
// det J (the recomputation of flagged points running with the accumulators dead) and the facet
// normals share the stage; each field's rows of the block leave as one contiguous run. K and the
// facet normals are the register kernels' formulas (jacobian.hpp, contraction off).
template <typename T, typename TN, bool KF>
__device__ __forceinline__ void finish_block(const typename Acc<T>::type (&acc)[4][3],
                                             const T (&x0)[3], const TN* __restrict__ nodes,
                                             const uint32_t* __restrict__ row, uint32_t k,
                                             const double* __restrict__ tab64,
                                             const T* __restrict__ ref_normals,
                                             const FactorParams<T>& p, uint64_t t, uint32_t qb,
                                             T* __restrict__ stage, const FactorOutputs& o,
                                             uint32_t lane) {
  constexpr bool kShift = sizeof(T) == 4;
  T* __restrict__ sj = stage;
  T* __restrict__ sx = stage;
  T* __restrict__ sd = stage + 3 * 256;
  T* __restrict__ sn = stage + 4 * 256;
  const uint32_t cell = lane & 15u, kg = lane >> 4;
  const uint32_t mask = p.mask;
  const uint32_t nv = p.nq - qb * 16u < 16u ? p.nq - qb * 16u : 16u;  // rows in the block
  // Register i of the product holds row 4 kg + i (FP32) or kg + 4 i (FP64) of the block.
  const auto qrow = [kg](int i) -> uint32_t {
    return sizeof(T) == 4 ? kg * 4u + static_cast<uint32_t>(i) : kg + 4u * static_cast<uint32_t>(i);
  };
  const auto jac_of = [&acc](int i, T (&j)[9]) {  // column-major: J(g, d) at d * 3 + g
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int g = 0; g < 3; ++g) j[d * 3 + g] = acc[1 + d][g][i];
  };
  T det[4];
  uint32_t near = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    T j[9];
    jac_of(i, j);
    det[i] = jac::det_of<T, 3, 3>(j);
    if constexpr (sizeof(T) == 4) {
      if (p.fixup) {
        float n2 = 0.f;
#pragma unroll
        for (int e = 0; e < 9; ++e) n2 = __builtin_fmaf(j[e], j[e], n2);
        if (__builtin_fabsf(det[i]) < static_cast<float>(p.tau) * n2 * __builtin_sqrtf(n2))
          near |= 1u << i;
      }
    }
  }
  const uint64_t r0 = t * p.nq + qb * 16u;  // the block's first row of the tile
  if (mask & kFJac) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      T j[9];
      jac_of(i, j);
#pragma unroll
      for (int e = 0; e < 9; ++e) sj[(qrow(i) * 9 + e) * 16 + cell] = j[e];
    }
    sf::wave_sync();
    flush(sj, gp(static_cast<T*>(o.j)) + ((r0 * 9) << 4), nv * 144u, lane);
    sf::wave_sync();  // J is read; the stage is free
  }
  T fn[4][3];
  if (KF && (mask & (kFPinv | kFFacetNormals))) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma clang fp contract(off)
      T j[9], kk[9];
      jac_of(i, j);
      jac::pinv_of<T, 3, 3>(j, kk);  // K(t, g) at g * 3 + t
      if (mask & kFPinv) {
#pragma unroll
        for (int e = 0; e < 9; ++e) sj[(qrow(i) * 9 + e) * 16 + cell] = kk[e];
      }
      if (mask & kFFacetNormals) {
        const uint32_t q = qb * 16u + qrow(i) < p.nq ? qb * 16u + qrow(i) : p.nq - 1;
        const T* __restrict__ rn = ref_normals + q * 3;
        T s2 = T(0);
#pragma unroll
        for (int g = 0; g < 3; ++g) {
          T a = kk[g * 3] * gp(rn)[0];
#pragma unroll
          for (int tt = 1; tt < 3; ++tt) a += kk[g * 3 + tt] * gp(rn)[tt];
          fn[i][g] = a;
          s2 = g == 0 ? a * a : s2 + a * a;
        }
        const T nrm = jac::jsqrt(s2);
#pragma unroll
        for (int g = 0; g < 3; ++g) fn[i][g] = fn[i][g] / nrm;
      }
    }
    if (mask & kFPinv) {
      sf::wave_sync();
      flush(sj, gp(static_cast<T*>(o.pinv)) + ((r0 * 9) << 4), nv * 144u, lane);
      sf::wave_sync();  // K is read; the stage is free
    }
  }
  if (mask & kFPoints) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
      for (int g = 0; g < 3; ++g)
        sx[(qrow(i) * 3 + g) * 16 + cell] =
            kShift ? (x0[g] + acc[0][g][i]) + p.origin[g] : acc[0][g][i] + p.origin[g];
    }
  }
  if (mask & kFDetJ) {
    if (near) {  // rare
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (near & (1u << i))
          det[i] = static_cast<T>(
              detj_fp64_k(nodes, row, k, p.last_node, tab64, qb * 16u + qrow(i)));
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) sd[qrow(i) * 16 + cell] = det[i];
  }
  if (KF && (mask & kFFacetNormals)) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int g = 0; g < 3; ++g) sn[(qrow(i) * 3 + g) * 16 + cell] = fn[i][g];
  }
  sf::wave_sync();
  if (mask & kFPoints) flush(sx, gp(static_cast<T*>(o.x)) + ((r0 * 3) << 4), nv * 48u, lane);
  if (mask & kFDetJ) flush(sd, gp(static_cast<T*>(o.detj)) + (r0 << 4), nv * 16u, lane);
  if (KF && (mask & kFFacetNormals))
    flush(sn, gp(static_cast<T*>(o.fnormals)) + ((r0 * 3) << 4), nv * 48u, lane);
  sf::wave_sync();  // the stage is read; the next block may overwrite it
}

// KF: the form that also forms K and the facet normals (their registers would cost the other
// calls a wavefront per SIMD). SCHED, the order of a tile's loads: 0, a table load per step and
// kind; 1, the table loads in groups ahead of the products; 2, as 1, and the next tile's dofmap
// entries loaded during this tile's products and its coordinates during this tile's output (more
// registers: fewer wavefronts per SIMD).
template <int K, typename TN, bool KF, int SCHED>
__global__ void __launch_bounds__(64 * kWaves)
    dense_mfma(const TN* __restrict__ nodes, const uint32_t* __restrict__ dofmap,
               const float* __restrict__ taba, const double* __restrict__ tab64,
               const float* __restrict__ ref_normals, FactorParams<float> p, FactorOutputs o) {
  constexpr int S = (K + 3) / 4;  // node steps of four (and gather rounds of four lanes a cell)
  __shared__ __align__(16) float lds[kWaves * kStageElems];
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t nqb = (p.nq + 15u) >> 4;
  float* __restrict__ stage = lds + w * kStageElems;

  // Tiles: XCD x = blockIdx.x % 8 takes the contiguous range [lo, hi) of the block's tiles; its
  // grid / 8 workgroups take four consecutive tiles (one per wavefront) at a time, in turn.
  const uint64_t ntiles = p.padded_cells >> 4;
  const uint32_t nx = gridDim.x >> 3, xcd = blockIdx.x & 7u, j = blockIdx.x >> 3;
  const uint64_t per = ntiles >> 3, rem = ntiles & 7u;
  const uint64_t lo = xcd * per + (xcd < rem ? xcd : rem);
  const uint64_t hi = lo + per + (xcd < rem ? 1u : 0u);

  const uint32_t cell = lane & 15u, kg = lane >> 4;
  const uint32_t mask = p.mask;
  const uint64_t step = static_cast<uint64_t>(nx) * kWaves;
  uint64_t t = lo + (static_cast<uint64_t>(j) * kWaves + w);
  if (t >= hi || nqb == 0) return;
  const auto row_of = [&](uint64_t tt) -> const uint32_t* {
    const uint64_t c0 = tt * 16 + cell;
    return dofmap + (c0 < p.num_cells ? c0 : p.num_cells - 1) * K;
  };
  // Gather: lane (cell, kg) loads node 0 and nodes kg, kg + 4, ... of its cell (past k: node
  // k - 1 again, so the loads go together). The dofmap entries (rr[0]: node 0, rr[1 + s]: node
  // 4 s + kg) and then the coordinates of the next tile are loaded while this tile's products
  // and output run, so a tile starts with its operands in registers.
  uint32_t rr[S + 1];
  TN x0n[3], v[S][3];
  const auto rows = [&](uint64_t tt) {
    const uint32_t* __restrict__ row = row_of(tt);
    rr[0] = gp(row)[0];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const uint32_t k = 4u * s + kg;
      rr[1 + s] = gp(row)[k < static_cast<uint32_t>(K) ? k : K - 1];
    }
  };
  const auto coords = [&]() {
    const uint32_t n0 = rr[0] < p.last_node ? rr[0] : p.last_node;
#pragma unroll
    for (int g = 0; g < 3; ++g) x0n[g] = gp(nodes)[static_cast<uint64_t>(n0) * 3 + g];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const uint32_t n = rr[1 + s] < p.last_node ? rr[1 + s] : p.last_node;
#pragma unroll
      for (int g = 0; g < 3; ++g) v[s][g] = gp(nodes)[static_cast<uint64_t>(n) * 3 + g];
    }
  };
  // B operands, formed in registers: the node lane (cell, kg) loads at step s is node 4 s + kg,
  // whose offset from node 0 (taken in the nodes' type, then rounded to FP32) is the operand
  // (zero for node 0 and past k).
  const auto offsets = [&](float (&x0)[3], float (&b)[S][3]) {
#pragma unroll
    for (int g = 0; g < 3; ++g) x0[g] = static_cast<float>(x0n[g]);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const uint32_t k = 4u * s + kg;
#pragma unroll
      for (int g = 0; g < 3; ++g)
        b[s][g] = (k > 0 && k < static_cast<uint32_t>(K)) ? static_cast<float>(v[s][g] - x0n[g])
                                                          : 0.f;
    }
  };
  const auto block_products = [&](uint32_t qb, const float (&b)[S][3], v4f (&acc)[4][3]) {
    const v4f zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int vv = 0; vv < 4; ++vv)
#pragma unroll
      for (int g = 0; g < 3; ++g) acc[vv][g] = zero;
    products<(SCHED > 0)>(mask, gp(taba) + qb * (4u * S * 64u) + lane, S, 0u, b, acc);
  };
  if constexpr (SCHED < 2) {
    for (; t < hi; t += step) {
      const uint32_t* __restrict__ row = row_of(t);
      rows(t);
      coords();
      float x0[3], b[S][3];
      offsets(x0, b);
      for (uint32_t qb = 0; qb < nqb; ++qb) {
        v4f acc[4][3];
        block_products(qb, b, acc);
        finish_block<float, TN, KF>(acc, x0, nodes, row, K, tab64, ref_normals, p, t, qb, stage,
                                    o, lane);
      }
    }
    return;
  }
  rows(t);
  coords();
  for (; t < hi; t += step) {
    const uint32_t* __restrict__ row = row_of(t);
    float x0[3], b[S][3];
    offsets(x0, b);
    // The last tile fetches its own entries again (in bounds, unused).
    rows(t + step < hi ? t + step : t);
    v4f acc[4][3];
    for (uint32_t qb = 0; qb + 1 < nqb; ++qb) {
      block_products(qb, b, acc);
      finish_block<float, TN, KF>(acc, x0, nodes, row, K, tab64, ref_normals, p, t, qb, stage, o,
                                  lane);
    }
    block_products(nqb - 1, b, acc);
    coords();  // the next tile's; b is dead from here
    finish_block<float, TN, KF>(acc, x0, nodes, row, K, tab64, ref_normals, p, t, nqb - 1, stage,
                                o, lane);
  }
}

// Cells of more nodes than the register operands above hold (runtime k), and every FP64 cell the
// matrix cores take: the same products, with the B operands gathered in chunks of steps for each
// block of 16 points, the accumulators carried across the chunks (the same node order). The work
// items are (tile, block of 16 points), so that the few tiles of high-degree meshes spread over
// more wavefronts; the XCDs take contiguous ranges of them, a tile's blocks in turn. In FP64
// (v_mfma_f64_16x16x4f64) the operands are the absolute coordinates from node 0, as the FP64
// kernels sum them; its stage of doubles takes two wavefronts a workgroup.
template <typename T>
constexpr int rt_waves() {
  return sizeof(T) == 4 ? kWaves : 2;
}
template <typename T>
constexpr int rt_chunk() {
  return sizeof(T) == 4 ? 8 : 4;  // steps of four nodes per chunk of B operands
}

template <typename T, typename TN, bool KF>
__global__ void __launch_bounds__(64 * rt_waves<T>())
    dense_mfma_rt(const TN* __restrict__ nodes, const uint32_t* __restrict__ dofmap,
                  const T* __restrict__ taba, const double* __restrict__ tab64,
                  const T* __restrict__ ref_normals, FactorParams<T> p, FactorOutputs o) {
  constexpr int kW = rt_waves<T>(), kC = rt_chunk<T>();
  constexpr bool kShift = sizeof(T) == 4;
  typedef typename Acc<T>::type V;
  __shared__ __align__(16) T lds[kW * kStageElems];
  const uint32_t K = p.k, S = (K + 3) / 4;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t nqb = (p.nq + 15u) >> 4;
  T* __restrict__ stage = lds + w * kStageElems;
  const uint64_t nitems = (p.padded_cells >> 4) * nqb;
  const uint32_t nx = gridDim.x >> 3, xcd = blockIdx.x & 7u, j = blockIdx.x >> 3;
  const uint64_t per = nitems >> 3, rem = nitems & 7u;
  const uint64_t lo = xcd * per + (xcd < rem ? xcd : rem);
  const uint64_t hi = lo + per + (xcd < rem ? 1u : 0u);
  const uint32_t cell = lane & 15u, kg = lane >> 4;
  const uint32_t mask = p.mask;
  for (uint64_t it = lo + (static_cast<uint64_t>(j) * kW + w); it < hi;
       it += static_cast<uint64_t>(nx) * kW) {
    const uint64_t t = it / nqb;
    const uint32_t qb = static_cast<uint32_t>(it - t * nqb);
    const uint64_t c0 = t * 16 + cell;
    const uint64_t c = c0 < p.num_cells ? c0 : p.num_cells - 1;
    const uint32_t* __restrict__ row = dofmap + c * K;
    TN x0n[3];
    T x0[3];
    {
      const uint32_t r0 = gp(row)[0];
      const uint32_t n0 = r0 < p.last_node ? r0 : p.last_node;
#pragma unroll
      for (int g = 0; g < 3; ++g) {
        x0n[g] = gp(nodes)[static_cast<uint64_t>(n0) * 3 + g];
        x0[g] = static_cast<T>(x0n[g]);
      }
    }
    const V zero = {0, 0, 0, 0};
    V acc[4][3];
#pragma unroll
    for (int vv = 0; vv < 4; ++vv)
#pragma unroll
      for (int g = 0; g < 3; ++g) acc[vv][g] = zero;
    gptr<const T> ta = gp(taba) + qb * (4u * S * 64u) + lane;
    for (uint32_t s0 = 0; s0 < S; s0 += kC) {
      TN v[kC][3];
#pragma unroll
      for (int i = 0; i < kC; ++i) {
        const uint32_t k = 4u * (s0 + i) + kg;
        const uint32_t r = gp(row)[k < K ? k : K - 1];
        const uint32_t n = r < p.last_node ? r : p.last_node;
#pragma unroll
        for (int g = 0; g < 3; ++g) v[i][g] = gp(nodes)[static_cast<uint64_t>(n) * 3 + g];
      }
      T b[kC][3];
#pragma unroll
      for (int i = 0; i < kC; ++i) {
        const uint32_t k = 4u * (s0 + i) + kg;
#pragma unroll
        for (int g = 0; g < 3; ++g) {
          if constexpr (kShift)
            b[i][g] = (k > 0 && k < K) ? static_cast<T>(v[i][g] - x0n[g]) : T(0);
          else
            b[i][g] = k < K ? static_cast<T>(v[i][g]) : T(0);
        }
      }
      products<false>(mask, ta, S, s0, b, acc);
    }
    finish_block<T, TN, KF>(acc, x0, nodes, row, K, tab64, ref_normals, p, t, qb, stage, o, lane);
  }
}

template <typename T, typename TN>
using DenseFn = void (*)(const TN*, const uint32_t*, const T*, const double*, const T*,
                         FactorParams<T>, FactorOutputs);

// 3D cells of 20 to 84 nodes. Measured on MI355X against the kernels they replace (X + J + det
// J, FP32): P3 tetrahedra 1.25x (the register kernel), Q2 hexahedra 1.11x and Q3 hexahedra 2.9x
// (the register kernels, where the sum-factorised ones do not apply), P3 pyramids 10x, P4
// tetrahedra 12x, P3 prisms 51x (the runtime-K kernel), P5 and P6 tetrahedra 14x and 9x, P4
// prisms 7.7x (the cooperative kernel). P2 tetrahedra (10 nodes), P2 pyramids (14) and P2
// prisms (18) measured slower here than on their register kernels and keep them.
// The load schedule of each node count (SCHED above), from X + J + det J in FP32 on MI355X
// against the table load per step (us per launch, one host): batched table loads for 56 nodes
// and more (P5 tetrahedra 82.2 -> 78.9, Q3 hexahedra at 100 points 1055 -> 833, P4 prisms 99.3
// -> 65.2, P6 tetrahedra 74.0 -> 61.6); with the prefetch for P3 tetrahedra (14 points 120.0 ->
// 112.9; 64 points 410 against 413) and P3 prisms (147.3 -> 125.3), in the plain form on FP32
// nodes only (the prefetch's registers); the table load per step for 27, 30 and 35 nodes (Q2
// hexahedra at 40 points 272 against 343 batched).
template <int K, typename TN, bool KF>
constexpr int sched_of() {
  constexpr int s = K >= 56 ? 1 : (K == 20 || K == 40) ? 2 : 0;
  return s == 2 && (KF || sizeof(TN) != 4) ? 1 : s;
}

template <typename TN, bool KF>
DenseFn<float, TN> pick(uint32_t k) {
  switch (k) {
    case 20:  // P3 tetrahedra
      return dense_mfma<20, TN, KF, sched_of<20, TN, KF>()>;
    case 27:  // Q2 hexahedra (point sets the sum-factorised kernels do not take)
      return dense_mfma<27, TN, KF, sched_of<27, TN, KF>()>;
    case 30:  // P3 pyramids
      return dense_mfma<30, TN, KF, sched_of<30, TN, KF>()>;
    case 35:  // P4 tetrahedra
      return dense_mfma<35, TN, KF, sched_of<35, TN, KF>()>;
    case 40:  // P3 prisms
      return dense_mfma<40, TN, KF, sched_of<40, TN, KF>()>;
    case 56:  // P5 tetrahedra
      return dense_mfma<56, TN, KF, sched_of<56, TN, KF>()>;
    case 64:  // Q3 hexahedra (as for Q2)
      return dense_mfma<64, TN, KF, sched_of<64, TN, KF>()>;
    case 75:  // P4 prisms
      return dense_mfma<75, TN, KF, sched_of<75, TN, KF>()>;
    case 84:  // P6 tetrahedra
      return dense_mfma<84, TN, KF, sched_of<84, TN, KF>()>;
    default:
      return nullptr;
  }
}

// FP32 node counts from 30 not listed above (P4 pyramids and the cells of more than 84 nodes)
// take the runtime-k form, as FP64 cells of 20 nodes and more do.
constexpr uint32_t kRtMinNodes = 30;
constexpr uint32_t kF64MinNodes = 20;

template <typename T, typename TN>
hipError_t launch(uint32_t k, const TN* nodes, const uint32_t* dofmap, const T* taba,
                  const double* tab64, const T* ref_normals, const FactorParams<T>& p,
                  const FactorOutputs& out, hipStream_t stream) {
  const bool kf = (p.mask & (kFPinv | kFFacetNormals)) != 0;
  DenseFn<T, TN> fn = nullptr;
  if constexpr (sizeof(T) == 4) fn = kf ? pick<TN, true>(k) : pick<TN, false>(k);
  const bool rt = fn == nullptr && k >= (sizeof(T) == 4 ? kRtMinNodes : kF64MinNodes);
  if (rt) fn = kf ? dense_mfma_rt<T, TN, true> : dense_mfma_rt<T, TN, false>;
  const int waves = rt ? rt_waves<T>() : kWaves;
  if (!fn || p.tile_log2 != 4 || p.nq == 0 || p.nq > kDenseMfmaMaxPoints || p.k != k ||
      ((p.mask & kFFacetNormals) && ref_normals == nullptr))
    return hipErrorNotSupported;
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  // The workgroups resident at once (registers and LDS), a multiple of 8, no more than the
  // tiles need.
  int per_cu = 0;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(fn),
                                                   64 * waves, 0);
  if (e != hipSuccess) return e;
  // Tiles of more than three blocks of points are bound by their write stream, which a
  // narrower band of tiles in flight writes faster: three workgroups a CU there (measured with
  // the table load per step, 116 VGPRs: P3 tetrahedra at 64 points 413.8 against 425.5 us with
  // four; at 14 points four measured faster, 120.8 against 125.0 us, and Q2 hexahedra at 40
  // points 272.6 against 285.1 us).
  if (!rt && (p.nq + 15) / 16 > 3 && per_cu > 3) per_cu = 3;
  // Work items: tiles, or (tile, block of 16 points) for the runtime-k form.
  const uint64_t items = (p.padded_cells >> 4) * (rt ? (p.nq + 15) / 16 : 1);
  uint64_t grid = static_cast<uint64_t>(cus > 0 ? cus : 8) * static_cast<uint64_t>(per_cu > 0 ? per_cu : 1);
  const uint64_t need = (items + waves - 1) / waves;
  const uint64_t need8 = (need + 7) / 8 * 8;
  grid = (grid < need8 ? grid : need8) / 8 * 8;
  if (grid < 8) grid = 8;
  hipLaunchKernelGGL(fn, dim3(static_cast<uint32_t>(grid)), dim3(64 * waves), 0, stream, nodes,
                     dofmap, taba, tab64, ref_normals, p, out);
  return hipGetLastError();
}

}  // namespace

bool has_dense_mfma(int tdim, int gdim, uint32_t k) {
  return tdim == 3 && gdim == 3 && (pick<float, false>(k) != nullptr || k >= kRtMinNodes);
}

bool has_dense_mfma_fp64(int tdim, int gdim, uint32_t k) {
  return tdim == 3 && gdim == 3 && k >= kF64MinNodes;
}

size_t dense_mfma_table_elems(uint32_t k, uint32_t nq) {
  const uint32_t S = (k + 3) / 4, nqb = (nq + 15) / 16;
  return static_cast<size_t>(nqb) * 4 * S * 64;
}

hipError_t launch_dense_mfma(uint32_t k, const float* nodes, const uint32_t* dofmap,
                             const float* taba, const double* tab64, const float* ref_normals,
                             const FactorParams<float>& p, const FactorOutputs& out,
                             hipStream_t stream) {
  return launch<float, float>(k, nodes, dofmap, taba, tab64, ref_normals, p, out, stream);
}

hipError_t launch_dense_mfma_f64_nodes(uint32_t k, const double* nodes, const uint32_t* dofmap,
                                       const float* taba, const double* tab64,
                                       const float* ref_normals, const FactorParams<float>& p,
                                       const FactorOutputs& out, hipStream_t stream) {
  return launch<float, double>(k, nodes, dofmap, taba, tab64, ref_normals, p, out, stream);
}

hipError_t launch_dense_mfma_fp64(uint32_t k, const double* nodes, const uint32_t* dofmap,
                                  const double* taba, const double* tab64,
                                  const double* ref_normals, const FactorParams<double>& p,
                                  const FactorOutputs& out, hipStream_t stream) {
  return launch<double, double>(k, nodes, dofmap, taba, tab64, ref_normals, p, out, stream);
}

}  // namespace tg
