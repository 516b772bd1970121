// Sum-factorised X, J and det J of Lagrange hexahedra at tensor-product point sets (the
// reference's `sumfact` strategy, tetraloma-geometry/src/sumfact/tensor.rs:176-337, as a
// schedule of the same arithmetic: X_g = sum_k x_kg phi_k, J_gd = sum_k x_kg d_d phi_k,
// layout.rs:138-164, with the sums taken one direction at a time).
//
// A hexahedron of degree n1 - 1 has n1^3 nodes (a, b, c) at (t_a, t_b, t_c) and its basis is
// the product of 1D Lagrange polynomials; at points (u_p, u_q, u_r) every sum over the n1^3
// nodes splits into three sums of n1 terms:
//   s_v(a, b)    = sum_c x(a, b, c) phi^(v)_c(u_r)          v = 0 (value), 1 (derivative)
//   S(a, q)_w    = sum_b s_v(a, b) phi^(u)_b(u_q)           w = (u, v) in {00, 10, 01}
//   X, J_xi, J_eta, J_zeta at (p, q, r) = sum_a phi^(t)_a(u_p) S(a, q)_w
// which is 14-fold less work than the dense n1^3 x m^3 contraction for Q3 at 4^3 points.
//
// One workgroup of two wavefronts per tile of 16 cells. They gather the tile's nodes through
// the dofmap (coalesced index loads, indices clamped to the last node) into LDS as
// [g][lex node][16 cells], so that the reads of the sums are conflict-free and the lanes of one
// cell share them. Lanes are 16 cells x 4 values of r, and each wavefront takes half of the
// values of q; each lane produces its points (p, q) of its r, one physical
// coordinate g at a time, and stores X_g and row g of J from registers. det J is built across
// the three rows (row0 x row1, then row2 . that), so no lane holds the whole 3 x 3 J of all its
// points at once. FP32: the coordinates are shifted to the mean of the cell's eight vertices,
// which halves the magnitude of the offsets against a shift to node 0; points where
// |det J| < tau |J|_F^3 get det J recomputed in FP64 from the stored coordinates.
#include <hip/hip_runtime.h>

#include "geom_internal.hpp"

namespace tg {
namespace {

constexpr int kBlock = 128;  // two wavefronts per tile of 16 cells

__device__ __forceinline__ float tfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double tfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float tsqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double tsqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float tabs(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ double tabs(double x) { return __builtin_fabs(x); }

template <typename T>
__device__ __forceinline__ void st(void* base, uint64_t idx, T v) {
  __builtin_nontemporal_store(v, reinterpret_cast<T*>(base) + idx);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// FP64 det J at point q of the lane's cell from its staged (stored-precision) coordinates: the
// dense sum over the K nodes with the f64 table, offsets from node 0 taken in FP64.
template <typename T, int K>
__device__ __noinline__ double detj_fp64_staged(const T* __restrict__ xs,
                                                const uint32_t* __restrict__ lex, uint32_t cell,
                                                const double* __restrict__ tab64, uint32_t q) {
  double x0[3], j[9];
  const uint32_t l0 = lex[0];
#pragma unroll
  for (int g = 0; g < 3; ++g) x0[g] = static_cast<double>(xs[(g * K + l0) * 16 + cell]);
#pragma unroll
  for (int i = 0; i < 9; ++i) j[i] = 0.0;
  for (int k = 1; k < K; ++k) {
    const uint32_t l = lex[k];
    const double* t = tab64 + (static_cast<uint64_t>(q) * K + k) * 4;
#pragma unroll
    for (int g = 0; g < 3; ++g) {
      const double d = static_cast<double>(xs[(g * K + l) * 16 + cell]) - x0[g];
#pragma unroll
      for (int dd = 0; dd < 3; ++dd) j[dd * 3 + g] = __builtin_fma(d, t[1 + dd], j[dd * 3 + g]);
    }
  }
  return j[0] * (j[4] * j[8] - j[7] * j[5]) - j[3] * (j[1] * j[8] - j[7] * j[2]) +
         j[6] * (j[1] * j[5] - j[4] * j[2]);
}

// N1 nodes and M points per direction. The two wavefronts of a workgroup share one tile: they
// gather it together, then wavefront h takes the points q in [h QH, (h + 1) QH) of the eta
// direction (QH = ceil(M / 2)), so that each lane keeps the sums of half the points. PC points of
// the p direction per pass (PC < M recomputes the first two sums to save registers, for FP64).
template <typename T, int N1, int M, int PC, bool SHIFT>
__global__ void __launch_bounds__(kBlock)
    factor_sumfact(const T* __restrict__ nodes, const uint32_t* __restrict__ dofmap,
                   const T* __restrict__ tab1d, const uint32_t* __restrict__ lex,
                   const double* __restrict__ tab64, FactorParams<T> p, FactorOutputs o) {
  constexpr int K = N1 * N1 * N1;
  constexpr int NQ = M * M * M;
  constexpr int QH = (M + 1) / 2;
  __shared__ T xs[3 * K * 16];
  const uint32_t t = threadIdx.x, cell = t & 15u, r = (t >> 4) & 3u, h = t >> 6;
  const uint64_t tile = blockIdx.x;

  // Gather: pair j = (cell, node) of the tile; the tile's dofmap rows are contiguous.
  for (uint32_t j = t; j < 16u * K; j += kBlock) {
    const uint32_t cc = j / K, node = j - cc * K;
    uint64_t c = tile * 16 + cc;
    c = c < p.num_cells ? c : p.num_cells - 1;
    uint32_t n = dofmap[c * K + node];
    n = n < p.last_node ? n : p.last_node;
    const uint32_t l = lex[node];
#pragma unroll
    for (int g = 0; g < 3; ++g) xs[(g * K + l) * 16 + cc] = nodes[static_cast<uint64_t>(n) * 3 + g];
  }
  __syncthreads();
  const int q0 = static_cast<int>(h) * QH;
  if (r >= static_cast<uint32_t>(M) || q0 >= M) return;  // no barrier follows

  T s[3];
#pragma unroll
  for (int g = 0; g < 3; ++g) {
    if constexpr (SHIFT) {
      T acc = T(0);
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int l = (((v & 1) * (N1 - 1)) * N1 + ((v >> 1) & 1) * (N1 - 1)) * N1 + (v >> 2) * (N1 - 1);
        acc += xs[(g * K + l) * 16 + cell];
      }
      s[g] = acc * T(0.125);
    } else {
      s[g] = T(0);
    }
  }
  // The lane's r: values and derivatives of the 1D basis at u_r.
  T br[N1], dr[N1];
#pragma unroll
  for (int c = 0; c < N1; ++c) {
    br[c] = tab1d[r * N1 + c];
    dr[c] = tab1d[(M + r) * N1 + c];
  }
  const uint32_t mask = p.mask;
  // The lane's first point (p, q, r) = (0, q0, r) and its field bases (tile layout, tile 16).
  const uint32_t first_point = static_cast<uint32_t>(M * q0) + M * M * r;
  const uint64_t row0 = tile * NQ + first_point;
  T* __restrict__ xbase = static_cast<T*>(o.x) + ((row0 * 3) << 4) + cell;
  T* __restrict__ jbase = static_cast<T*>(o.j) + ((row0 * 9) << 4) + cell;
  T* __restrict__ dbase = static_cast<T*>(o.detj) + (row0 << 4) + cell;

#pragma unroll
  for (int p0 = 0; p0 < M; p0 += PC) {
    T carry[QH][PC][3], n2[QH][PC];
    // g and a stay rolled: unrolled, the compiler hoists the LDS reads of later iterations and
    // the registers they hold cost occupancy. Sums start from zero (fma(x, y, 0) = x y).
#pragma unroll 1
    for (int g = 0; g < 3; ++g) {
      const T sg = g == 0 ? s[0] : (g == 1 ? s[1] : s[2]);
      const T* __restrict__ xg = xs + g * K * 16 + cell;
      T acc[QH][PC][4];
#pragma unroll
      for (int qq = 0; qq < QH; ++qq)
#pragma unroll
        for (int pp = 0; pp < PC; ++pp)
#pragma unroll
          for (int i = 0; i < 4; ++i) acc[qq][pp][i] = T(0);
#pragma unroll 1
      for (int a = 0; a < N1; ++a) {
        T s2[QH][3];
#pragma unroll
        for (int qq = 0; qq < QH; ++qq) s2[qq][0] = s2[qq][1] = s2[qq][2] = T(0);
#pragma unroll
        for (int b = 0; b < N1; ++b) {
          T s0 = T(0), s1 = T(0);
#pragma unroll
          for (int c = 0; c < N1; ++c) {
            const T x = xg[((a * N1 + b) * N1 + c) * 16] - sg;
            s0 = tfma(x, br[c], s0);
            s1 = tfma(x, dr[c], s1);
          }
#pragma unroll
          for (int qq = 0; qq < QH; ++qq) {
            const int q = q0 + qq < M ? q0 + qq : M - 1;
            const T bq = tab1d[q * N1 + b], dq = tab1d[(M + q) * N1 + b];
            s2[qq][0] = tfma(s0, bq, s2[qq][0]);
            s2[qq][1] = tfma(s0, dq, s2[qq][1]);
            s2[qq][2] = tfma(s1, bq, s2[qq][2]);
          }
        }
#pragma unroll
        for (int qq = 0; qq < QH; ++qq)
#pragma unroll
          for (int pp = 0; pp < PC; ++pp) {
            if (p0 + pp >= M) continue;
            const T ba = tab1d[(p0 + pp) * N1 + a], da = tab1d[(M + p0 + pp) * N1 + a];
            T* A = acc[qq][pp];
            A[0] = tfma(s2[qq][0], ba, A[0]);
            A[1] = tfma(s2[qq][0], da, A[1]);
            A[2] = tfma(s2[qq][1], ba, A[2]);
            A[3] = tfma(s2[qq][2], ba, A[3]);
          }
      }
      // Row g of J and X_g of the lane's points; the det J carry. Addresses: one per-lane base
      // per field and compile-time offsets per point.
      const T og = g == 0 ? p.origin[0] : (g == 1 ? p.origin[1] : p.origin[2]);
      T* __restrict__ xb = xbase + g * 16;
      T* __restrict__ jb = jbase + g * 16;
#pragma unroll
      for (int qq = 0; qq < QH; ++qq)
#pragma unroll
        for (int pp = 0; pp < PC; ++pp) {
          if (p0 + pp >= M || q0 + qq >= M) continue;
          const int pt = p0 + pp + M * qq;  // point offset from the lane's first point
          const T* A = acc[qq][pp];
          if (mask & kFPoints) st<T>(xb, pt * 3 * 16, (A[0] + sg) + og);
          if (mask & kFJac) {
#pragma unroll
            for (int d = 0; d < 3; ++d) st<T>(jb, (pt * 9 + d * 3) * 16, A[1 + d]);
          }
          T* cr = carry[qq][pp];
          const T nrm = A[1] * A[1] + A[2] * A[2] + A[3] * A[3];
          if (g == 0) {
            cr[0] = A[1];
            cr[1] = A[2];
            cr[2] = A[3];
            n2[qq][pp] = nrm;
          } else if (g == 1) {  // row0 x row1
            const T c0 = cr[1] * A[3] - cr[2] * A[2];
            const T c1 = cr[2] * A[1] - cr[0] * A[3];
            const T c2 = cr[0] * A[2] - cr[1] * A[1];
            cr[0] = c0;
            cr[1] = c1;
            cr[2] = c2;
            n2[qq][pp] += nrm;
          } else {  // det J = row2 . (row0 x row1)
            cr[0] = A[1] * cr[0] + A[2] * cr[1] + A[3] * cr[2];
            n2[qq][pp] += nrm;
          }
        }
    }
    if (mask & kFDetJ) {
#pragma unroll
      for (int qq = 0; qq < QH; ++qq)
#pragma unroll
        for (int pp = 0; pp < PC; ++pp) {
          if (p0 + pp >= M || q0 + qq >= M) continue;
          const int pt = p0 + pp + M * qq;
          T det = carry[qq][pp][0];
          if constexpr (sizeof(T) == 4) {
            const T nn = n2[qq][pp];
            if (p.fixup && tabs(det) < static_cast<T>(p.tau) * nn * tsqrt(nn))
              det = static_cast<T>(
                  detj_fp64_staged<T, K>(xs, lex, cell, tab64, first_point + static_cast<uint32_t>(pt)));
          }
          st<T>(dbase, pt * 16, det);
        }
    }
  }
}

template <typename T>
using SumfactFn = void (*)(const T*, const uint32_t*, const T*, const uint32_t*, const double*,
                           FactorParams<T>, FactorOutputs);

// FP32 keeps its points of an r-slab in registers; FP64 takes the p direction two at a time.
template <typename T, int N1, int M, bool SHIFT>
constexpr SumfactFn<T> sumfact_fn() {
  constexpr int PC = sizeof(T) == 4 ? M : (M < 2 ? M : 2);  // FP64: p two at a time
  return factor_sumfact<T, N1, M, PC, SHIFT>;
}

template <typename T, bool SHIFT>
SumfactFn<T> pick(int n1, int m) {
#define TG_SF(N1_, M_) \
  if (n1 == N1_ && m == M_) return sumfact_fn<T, N1_, M_, SHIFT>();
  TG_SF(2, 1) TG_SF(2, 2) TG_SF(2, 3) TG_SF(2, 4)
  TG_SF(3, 1) TG_SF(3, 2) TG_SF(3, 3) TG_SF(3, 4)
  TG_SF(4, 1) TG_SF(4, 2) TG_SF(4, 3) TG_SF(4, 4)
#undef TG_SF
  return nullptr;
}

}  // namespace

template <typename T>
hipError_t launch_sumfact(int n1, int m, bool shift, const T* nodes, const uint32_t* dofmap,
                          const T* tab1d, const uint32_t* lex, const double* tab64,
                          const FactorParams<T>& p, const FactorOutputs& out, hipStream_t stream) {
  SumfactFn<T> fn = shift ? pick<T, true>(n1, m) : pick<T, false>(n1, m);
  if (!fn || p.tile_log2 != 4) return hipErrorNotSupported;
  const uint64_t blocks = p.padded_cells >> 4;  // one workgroup per tile
  if (blocks > 0x7fffffffu) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(fn, dim3(static_cast<uint32_t>(blocks)), dim3(kBlock), 0, stream, nodes,
                     dofmap, tab1d, lex, tab64, p, out);
  return hipGetLastError();
}

template hipError_t launch_sumfact<float>(int, int, bool, const float*, const uint32_t*,
                                          const float*, const uint32_t*, const double*,
                                          const FactorParams<float>&, const FactorOutputs&,
                                          hipStream_t);
template hipError_t launch_sumfact<double>(int, int, bool, const double*, const uint32_t*,
                                           const double*, const uint32_t*, const double*,
                                           const FactorParams<double>&, const FactorOutputs&,
                                           hipStream_t);

}  // namespace tg
