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
// The schedule (one 256-thread workgroup per tile of 16 cells) is described at sf::tile in
// sumfact_tile.hpp. FP32: the
// coordinates are shifted to the mean of each cell's eight vertices before the sums, which
// halves the offsets against a shift to node 0; points where |det J| < tau |J|_F^3 get det J
// recomputed in FP64 from the stored coordinates.
#include <hip/hip_runtime.h>

#include "geom_internal.hpp"
#include "sumfact_tile.hpp"

namespace tg {
namespace {

template <typename T, int N1, int M, int PC, bool SHIFT, typename TN = T, typename TO = T>
__global__ void __launch_bounds__(sf::kThreads)
    factor_sumfact(const TN* __restrict__ nodes, const uint32_t* __restrict__ dofmap,
                   const T* __restrict__ tab1d, const uint32_t* __restrict__ lex,
                   const double* __restrict__ tab64, FactorParams<TO> p, FactorOutputs o) {
  __shared__ __align__(16) T lds[sf::Lds<N1, M, PC>::kTotal];
  // Tiles in dispatch order: giving each XCD a contiguous range of tiles instead (as the
  // per-block dense kernels do) measured 9-14 % slower here.
  sf::tile<T, N1, M, PC, SHIFT, TN, TO>(nodes, dofmap, tab1d, lex, tab64, p, o, blockIdx.x, lds);
}

template <typename T, typename TN = T>
using SumfactFn = void (*)(const TN*, const uint32_t*, const T*, const uint32_t*, const double*,
                           FactorParams<T>, FactorOutputs);

// FP32 keeps its points of an r-slab in registers; FP64 takes the p direction two at a time.
template <typename T, int N1, int M, bool SHIFT, typename TN>
constexpr SumfactFn<T, TN> sumfact_fn() {
  constexpr int PC = sizeof(T) == 4 ? M : (M < 2 ? M : 2);  // FP64: p two at a time
  return factor_sumfact<T, N1, M, PC, SHIFT, TN>;
}

template <typename T, bool SHIFT, typename TN = T>
SumfactFn<T, TN> pick(int n1, int m) {
#define TG_SF(N1_, M_) \
  if (n1 == N1_ && m == M_) return sumfact_fn<T, N1_, M_, SHIFT, TN>();
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
  hipLaunchKernelGGL(fn, dim3(static_cast<uint32_t>(blocks)), dim3(sf::kThreads), 0, stream, nodes,
                     dofmap, tab1d, lex, tab64, p, out);
  return hipGetLastError();
}

template hipError_t launch_sumfact<float>(int, int, bool, const float*, const uint32_t*,
                                          const float*, const uint32_t*, const double*,
                                          const FactorParams<float>&, const FactorOutputs&,
                                          hipStream_t);
// FP64 nodes of an FP32 geometry: the sums in FP64 (as the FP64 kernel, unshifted, p two at a
// time) and X, J, det J rounded to FP32 at their stores.
using AccurateFn = void (*)(const double*, const uint32_t*, const double*, const uint32_t*,
                            const double*, FactorParams<float>, FactorOutputs);

AccurateFn pick_accurate(int n1, int m) {
#define TG_SFA(N1_, M_)        \
  if (n1 == N1_ && m == M_)    \
    return factor_sumfact<double, N1_, M_, (M_ < 2 ? M_ : 2), false, double, float>;
  TG_SFA(2, 1) TG_SFA(2, 2) TG_SFA(2, 3) TG_SFA(2, 4)
  TG_SFA(3, 1) TG_SFA(3, 2) TG_SFA(3, 3) TG_SFA(3, 4)
  TG_SFA(4, 1) TG_SFA(4, 2) TG_SFA(4, 3) TG_SFA(4, 4)
#undef TG_SFA
  return nullptr;
}

hipError_t launch_sumfact_f64_nodes(int n1, int m, const double* nodes, const uint32_t* dofmap,
                                    const double* tab1d, const uint32_t* lex, const double* tab64,
                                    const FactorParams<float>& p, const FactorOutputs& out,
                                    hipStream_t stream) {
  AccurateFn fn = pick_accurate(n1, m);
  if (!fn || p.tile_log2 != 4) return hipErrorNotSupported;
  const uint64_t blocks = p.padded_cells >> 4;
  if (blocks > 0x7fffffffu) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(fn, dim3(static_cast<uint32_t>(blocks)), dim3(sf::kThreads), 0, stream, nodes,
                     dofmap, tab1d, lex, tab64, p, out);
  return hipGetLastError();
}

template hipError_t launch_sumfact<double>(int, int, bool, const double*, const uint32_t*,
                                           const double*, const uint32_t*, const double*,
                                           const FactorParams<double>&, const FactorOutputs&,
                                           hipStream_t);

}  // namespace tg
