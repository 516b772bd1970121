// Sum-factorised X, J and det J of FP32 Lagrange hexahedra with the last two contractions on the
// matrix cores (v_mfma_f32_16x16x4f32). Same arithmetic as sumfact_kernels.hip (the reference's
// `sumfact` strategy, tetraloma-geometry/src/sumfact/tensor.rs:176-337, over the per-point
// formulas of layout.rs:138-213), scheduled for CDNA's 64-wide wavefronts and matrix cores:
//
//   s_v(a, b; r)       = sum_c x(a, b, c) phi^(v)_c(u_r)                 (vector ALU)
//   X    (p, q, r)     = sum_{a,b} [phi_a(u_p) phi_b(u_q)]   s_0(a, b; r)   (matrix cores)
//   J_xi (p, q, r)     = sum_{a,b} [phi'_a(u_p) phi_b(u_q)]  s_0(a, b; r)
//   J_eta(p, q, r)     = sum_{a,b} [phi_a(u_p) phi'_b(u_q)]  s_0(a, b; r)
//   J_zeta(p, q, r)    = sum_{a,b} [phi_a(u_p) phi_b(u_q)]   s_1(a, b; r)
//
// Lane l of a wavefront is (a = l / 16, cell = l % 16) of a tile of 16 cells. The lane holds the
// n1^2 nodes (a, *, *) of its cell and takes the first sums (along zeta) in registers. The
// second and third sums (along eta and xi together, K = n1^2 = (a, b)) are one matrix product
// per (r, coordinate, kind): D[(p, q)][cell] = sum_b sum_a T[(p, q)][(a, b)] s(a, b)[cell],
// issued as n1 v_mfma_f32_16x16x4f32 steps (one per b) whose B operand is exactly the lane's
// own register s(a = l / 16, b)[cell = l % 16] and whose A operand is the 2D table value of the
// lane (T[(p, q) = l % 16][a = l / 16] at that b, rounded once from FP64 on the host; staged in
// LDS once per workgroup). The product lands in lane (q = l / 16, cell = l % 16), register p:
// the lane holds X and J of the points (0..m-1, q, r) of its cell and forms det J (FP64
// recomputation where |det J| < tau |J|_F^3, as every FP32 kernel here).
//
// The schedule is set by the output stream, 85 % of the bytes: how fast HBM absorbs the writes
// depends on how scattered the addresses written at one time are. So each tile is one
// workgroup of four wavefronts, wavefront w computing the points with r = w: the points
// (0..m-1, 0..m-1, r) are consecutive in the point order, and the four wavefronts write the
// tile's X, J and det J as three contiguous runs (staged in LDS, 16-byte non-temporal stores).
// The workgroups are persistent and few per CU, and they walk the tiles interleaved (workgroup i
// takes tiles i, i + grid, ...), so the tiles being written at any time form one narrow band of
// the fields (on Q3 96^3 2.4 % faster than one contiguous range per XCD, 0.673 against
// 0.689 ms, though neighbouring tiles then meet their shared nodes in different L2s).
// With few wavefronts to hide latency, each workgroup pipelines its gather: wavefront w loads node
// slice b = w of a tile (the lane's nodes (a, w, *)) one tile ahead, and the dofmap entries of
// those nodes two tiles ahead, so both are in flight while a tile is computed and written; a
// tile's slices meet in LDS, from where every wavefront reads the whole tile.
//
// FP32: the coordinates are shifted to the mean of each cell's eight vertices before the sums
// (summed in vertex order, as sumfact_kernels.hip), and X adds it back.
//
// FP64-stored nodes of an FP32 geometry (TN = double): the gathered coordinates stay in FP64 in
// LDS; the vertex mean, the offsets from it and the first sums are taken in FP64, and each first
// sum s enters the matrix cores as a hi/lo pair, hi = fl32(s) and lo = fl32(s - hi): X takes the
// hi products, J the hi and then the lo products into the same FP32 accumulator. What remains in
// FP32 is the 2D table rounding and the accumulation, about 3e-7 of |J| on curved Q3 hexahedra.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <type_traits>

#include "geom_internal.hpp"
#include "sumfact_tile.hpp"

namespace tg {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;  // one tile per workgroup: wavefront w gathers slice b = w, computes r = w

__device__ __forceinline__ v4f mfma4(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Workgroup barrier for LDS only: waits for this wavefront's LDS operations, not for its global
// loads and stores in flight (the prefetches must stay in flight across it).
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Copies N values (a multiple of 4) from LDS to global memory with 16-byte non-temporal stores,
// lane by lane over the wavefront; the number of stores is fixed at compile time.
template <int N>
__device__ __forceinline__ void flush(const float* __restrict__ src, float* __restrict__ dst,
                                      uint32_t lane) {
  static_assert(N % 4 == 0, "whole 16-byte vectors");
  constexpr int V = N / 4;
#pragma unroll
  for (int i = 0; i < (V + 63) / 64; ++i) {
    const int e = i * 64 + static_cast<int>(lane);
    if (V % 64 == 0 || e < V) {
      const v4f v = *reinterpret_cast<const v4f*>(src + e * 4);
      __builtin_nontemporal_store(v, reinterpret_cast<v4f*>(dst + e * 4));
    }
  }
}

// LDS of one workgroup: the A operands, the four wavefronts' output stages (floats), then the
// tile's nodes (TN). FP64 nodes: 80,896 bytes for Q3, two workgroups per CU as for FP32.
template <int N1, int M, typename TN>
struct StreamLds {
  static constexpr int kTab = 3 * N1 * 64;         // the A operands [kind][b][lane]
  static constexpr int kX = 3 * N1 * N1 * 64;      // the tile's nodes [g][b][c][lane]
  static constexpr int kRun = M * M * 16;          // one component over the points of one r
  static constexpr int kStage = 13 * kRun;         // per wavefront: its X, J and det J runs
  static constexpr int kFloats = kTab + kWaves * kStage;
  static constexpr int kBytes = kFloats * 4 + kX * static_cast<int>(sizeof(TN));
  static_assert(kTab % 4 == 0 && kRun % 4 == 0, "16-byte aligned stages");
};

// FULL: the factors are X, J and det J (fixed at compile time, so that the loop is straight-line
// code and the waits on the prefetches count the stores exactly); otherwise any subset of them,
// from p.mask. TN: the stored node type (double: FP64 nodes, always shifted); tab1d is of TN.
template <int N1, int M, bool SHIFT, bool FULL, typename TN>
__global__ void __launch_bounds__(64 * kWaves)
    factor_sumfact_mfma(const TN* __restrict__ nodes, const uint32_t* __restrict__ dofmap,
                        const TN* __restrict__ tab1d, const uint32_t* __restrict__ alex,
                        const float* __restrict__ tab2d, const double* __restrict__ tab64,
                        FactorParams<float> p, FactorOutputs o) {
  constexpr int K = N1 * N1 * N1;
  constexpr int NQ = M * M * M;
  constexpr bool N64 = std::is_same<TN, double>::value;
  static_assert(!N64 || SHIFT, "FP64 nodes are always shifted");
  using L = StreamLds<N1, M, TN>;
  constexpr int kRun = L::kRun;
  static_assert(N1 >= 2 && N1 <= 4 && M >= 1 && M <= 4, "one 16x16x4 product per (b, kind)");
  __shared__ __align__(16) unsigned char smem[L::kBytes];
  float* __restrict__ lds = reinterpret_cast<float*>(smem);
  TN* __restrict__ xs = reinterpret_cast<TN*>(smem + L::kFloats * 4);
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) & (kWaves - 1);
  // The wavefront's stage: X [M^2 points][3][16], then J [M^2][9][16], then det J [M^2][16].
  float* __restrict__ stage = lds + L::kTab + w * L::kStage;
  float* __restrict__ sx = stage;
  float* __restrict__ sj = stage + 3 * kRun;
  float* __restrict__ sd = stage + 12 * kRun;
  for (int i = static_cast<int>(threadIdx.x); i < L::kTab; i += 64 * kWaves) lds[i] = tab2d[i];

  // Tiles: workgroup i takes tiles i, i + grid, i + 2 grid, ... of the block, so that the tiles
  // in flight on the whole chip form one band of the fields.
  const uint64_t hi = p.padded_cells >> 4;
  const uint32_t nx = gridDim.x;
  if (blockIdx.x >= hi) return;  // the whole workgroup: no tile

  const uint32_t cell = lane & 15u, a = lane >> 4;
  const bool a_live = a < static_cast<uint32_t>(N1);
  const uint32_t mask = FULL ? (kFPoints | kFJac | kFDetJ) : p.mask;
  // A.5 index of the lane's slice nodes (a, w, c); every lane loads (lanes and wavefronts past
  // n1 a valid node they then drop), so that the prefetches are branch-free.
  uint32_t map[N1];
  const uint32_t ga = a_live ? a : 0u, gw = w < static_cast<uint32_t>(N1) ? w : 0u;
#pragma unroll
  for (int cc = 0; cc < N1; ++cc) map[cc] = alex[(ga * N1 + gw) * N1 + cc];
  auto cell_of = [&](uint64_t t) {
    const uint64_t c = t * 16 + cell;
    return c < p.num_cells ? c : p.num_cells - 1;
  };
  auto load_idx = [&](uint64_t t, uint32_t (&idx)[N1]) {
    const uint32_t* __restrict__ row = dofmap + cell_of(t) * K;
#pragma unroll
    for (int cc = 0; cc < N1; ++cc) idx[cc] = row[map[cc]];
  };
  struct P3 {
    TN v[3];
  };
  auto load_xyz = [&](const uint32_t (&idx)[N1], P3 (&v)[N1]) {
#pragma unroll
    for (int cc = 0; cc < N1; ++cc) {
      const uint32_t n = idx[cc] < p.last_node ? idx[cc] : p.last_node;  // clamped indices
      v[cc] = *reinterpret_cast<const P3*>(nodes + static_cast<uint64_t>(n) * 3);
    }
  };
  auto last_or = [&](uint64_t t) { return t < hi ? t : hi - 1; };  // prefetch past the end: repeat

  const uint32_t q = a;  // the output lane's point row
  const bool q_live = q < static_cast<uint32_t>(M);
  const int r = static_cast<int>(w);
  float* __restrict__ ox = static_cast<float*>(o.x);
  float* __restrict__ oj = static_cast<float*>(o.j);
  float* __restrict__ od = static_cast<float*>(o.detj);

  // Tile t's nodes from the registers of its prefetch into xs (slice b = w).
  auto put_slice = [&](const P3 (&v)[N1]) {
    if (w < static_cast<uint32_t>(N1)) {
#pragma unroll
      for (int g = 0; g < 3; ++g)
#pragma unroll
        for (int cc = 0; cc < N1; ++cc)
          xs[((g * N1 + w) * N1 + cc) * 64 + lane] = a_live ? v[cc].v[g] : TN(0);
    }
  };
  // X, J and det J of tile t at the points with r = w, from xs into the wavefront's stage.
  auto compute = [&](uint64_t t) {
    const v4f zero = {0.f, 0.f, 0.f, 0.f};
    v4f ax[3], jac[3][3];  // jac[d][g] = J(g, d) at the lane's points (p = register index)
    // The products of coordinate g from its first sums: X from hi, J from hi, then from lo.
    auto products = [&](int g, const float (&h0)[N1], const float (&h1)[N1],
                        const float (&l0)[N1], const float (&l1)[N1]) {
      v4f px = zero, j0 = zero, j1 = zero, j2 = zero;
#pragma unroll
      for (int b = 0; b < N1; ++b) {
        const float t_val = lds[b * 64 + lane], t_xi = lds[(N1 + b) * 64 + lane],
                    t_eta = lds[(2 * N1 + b) * 64 + lane];
        if (mask & kFPoints) px = mfma4(t_val, h0[b], px);
        j0 = mfma4(t_xi, h0[b], j0);
        j1 = mfma4(t_eta, h0[b], j1);
        j2 = mfma4(t_val, h1[b], j2);
        if constexpr (N64) {
          j0 = mfma4(t_xi, l0[b], j0);
          j1 = mfma4(t_eta, l0[b], j1);
          j2 = mfma4(t_val, l1[b], j2);
        }
      }
      ax[g] = px;
      jac[0][g] = j0;
      jac[1][g] = j1;
      jac[2][g] = j2;
    };
    // The cell's vertex mean: vertex k is node (a, b, c) = (k & 1, k >> 1 & 1, k >> 2) (n1 - 1),
    // at lane ((k & 1) (n1 - 1)) 16 + cell of xs; summed from vertex 0 in order.
    auto vertex_mean = [&](int g) {
      constexpr int E = N1 - 1;
      TN acc = TN(0);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const TN vk = xs[((g * N1 + ((k >> 1) & 1) * E) * N1 + ((k >> 2) & 1) * E) * 64 +
                         (k & 1) * E * 16 + cell];
        acc = k == 0 ? vk : acc + vk;
      }
      return acc * TN(0.125);
    };
    TN sg[3] = {TN(0), TN(0), TN(0)};
    if constexpr (!N64) {
      float x[3][N1][N1];
#pragma unroll
      for (int g = 0; g < 3; ++g)
#pragma unroll
        for (int b = 0; b < N1; ++b)
#pragma unroll
          for (int cc = 0; cc < N1; ++cc) x[g][b][cc] = xs[((g * N1 + b) * N1 + cc) * 64 + lane];
      if constexpr (SHIFT) {
#pragma unroll
        for (int g = 0; g < 3; ++g) {
          sg[g] = vertex_mean(g);
#pragma unroll
          for (int b = 0; b < N1; ++b)
#pragma unroll
            for (int cc = 0; cc < N1; ++cc) x[g][b][cc] = a_live ? x[g][b][cc] - sg[g] : 0.f;
        }
      }
      // First sums (along zeta) for r, from the c = 0 product in ascending c, then the products.
#pragma unroll
      for (int g = 0; g < 3; ++g) {
        float s0[N1], s1[N1];
#pragma unroll
        for (int b = 0; b < N1; ++b) {
          float u0 = 0.f, u1 = 0.f;
#pragma unroll
          for (int cc = 0; cc < N1; ++cc) {
            u0 = __builtin_fmaf(x[g][b][cc], tab1d[r * N1 + cc], u0);
            u1 = __builtin_fmaf(x[g][b][cc], tab1d[(M + r) * N1 + cc], u1);
          }
          s0[b] = u0;
          s1[b] = u1;
        }
        products(g, s0, s1, s0, s1);
      }
    } else {
      // One coordinate at a time (16 FP64 values live): mean, offsets and first sums in FP64,
      // each sum split into hi and lo.
#pragma unroll
      for (int g = 0; g < 3; ++g) {
        sg[g] = vertex_mean(g);
        float h0[N1], h1[N1], l0[N1], l1[N1];
#pragma unroll
        for (int b = 0; b < N1; ++b) {
          double u0 = 0.0, u1 = 0.0;
#pragma unroll
          for (int cc = 0; cc < N1; ++cc) {
            const double d = a_live ? xs[((g * N1 + b) * N1 + cc) * 64 + lane] - sg[g] : 0.0;
            u0 = __builtin_fma(d, tab1d[r * N1 + cc], u0);
            u1 = __builtin_fma(d, tab1d[(M + r) * N1 + cc], u1);
          }
          h0[b] = static_cast<float>(u0);
          h1[b] = static_cast<float>(u1);
          l0[b] = static_cast<float>(u0 - static_cast<double>(h0[b]));
          l1[b] = static_cast<float>(u1 - static_cast<double>(h1[b]));
        }
        products(g, h0, h1, l0, l1);
      }
    }
    float det[M];
    uint32_t near = 0;  // bit p: point (p, q, r) needs the FP64 det J
#pragma unroll
    for (int pp = 0; pp < M; ++pp) {
      // First-row cofactor expansion of the column-major J (layout.rs:186-213).
      det[pp] = jac[0][0][pp] * (jac[1][1][pp] * jac[2][2][pp] - jac[2][1][pp] * jac[1][2][pp]) -
                jac[1][0][pp] * (jac[0][1][pp] * jac[2][2][pp] - jac[2][1][pp] * jac[0][2][pp]) +
                jac[2][0][pp] * (jac[0][1][pp] * jac[1][2][pp] - jac[1][1][pp] * jac[0][2][pp]);
      if (p.fixup) {
        float n2 = 0.f;
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
          for (int g = 0; g < 3; ++g) n2 = __builtin_fmaf(jac[d][g][pp], jac[d][g][pp], n2);
        if (__builtin_fabsf(det[pp]) < static_cast<float>(p.tau) * n2 * __builtin_sqrtf(n2))
          near |= 1u << pp;
      }
    }
    if (!q_live) return;
    if (mask & kFPoints)
#pragma unroll
      for (int pp = 0; pp < M; ++pp)
#pragma unroll
        for (int g = 0; g < 3; ++g)
          sx[((pp + M * q) * 3 + g) * 16 + cell] =
              N64 ? static_cast<float>((static_cast<double>(ax[g][pp]) + sg[g]) + p.origin[g])
                  : static_cast<float>((ax[g][pp] + sg[g]) + p.origin[g]);
    if (mask & kFJac)
#pragma unroll
      for (int pp = 0; pp < M; ++pp)
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
          for (int g = 0; g < 3; ++g)
            sj[((pp + M * q) * 9 + d * 3 + g) * 16 + cell] = jac[d][g][pp];
    if (mask & kFDetJ) {
      if (near) {  // rare: after the products, so that the call keeps few values live
        const uint32_t* __restrict__ row = dofmap + cell_of(t) * K;
#pragma unroll
        for (int pp = 0; pp < M; ++pp)
          if (near & (1u << pp))
            det[pp] = static_cast<float>(sf::detj_fp64<K>(
                nodes, row, p.last_node, tab64, static_cast<uint32_t>(pp + M * q + M * M * r)));
      }
#pragma unroll
      for (int pp = 0; pp < M; ++pp) sd[(pp + M * q) * 16 + cell] = det[pp];
    }
  };
  // The wavefront's runs of tile t (the points (0..m-1, 0..m-1, r)) from its stage.
  auto flush_tile = [&](uint64_t t) {
    sf::wave_sync();
    const uint64_t run0 = t * NQ + static_cast<uint32_t>(M * M * r);
    if (mask & kFPoints) flush<3 * kRun>(sx, ox + ((run0 * 3) << 4), lane);
    if (mask & kFJac) flush<9 * kRun>(sj, oj + ((run0 * 9) << 4), lane);
    if (mask & kFDetJ) flush<kRun>(sd, od + (run0 << 4), lane);
  };

  // Tile t0 (peeled, so that the loop body is straight-line): its nodes, and in flight the
  // nodes of t0 + nx and the entries of t0 + 2 nx.
  uint64_t t = blockIdx.x;
  uint32_t idx[N1];
  P3 v[N1];
  load_idx(t, idx);
  load_xyz(idx, v);
  load_idx(last_or(t + nx), idx);
  lds_barrier();  // the tables
  put_slice(v);
  lds_barrier();
  load_xyz(idx, v);
  load_idx(last_or(t + 2 * nx), idx);
  if (r < M) compute(t);
#pragma unroll 1
  for (t += nx; t < hi; t += nx) {
    lds_barrier();  // every wavefront is done with the previous tile's nodes
    put_slice(v);   // waits for tile t's nodes (issued a tile ago)
    lds_barrier();
    if (r < M) flush_tile(t - nx);  // the previous tile leaves while this one is computed
    load_xyz(idx, v);
    load_idx(last_or(t + 2 * nx), idx);
    if (r < M) {
      sf::wave_sync();  // the stage is free again once its reads above have returned
      compute(t);
    }
  }
  if (r < M) flush_tile(t - nx);
}

template <typename TN>
using MfmaFn = void (*)(const TN*, const uint32_t*, const TN*, const uint32_t*, const float*,
                        const double*, FactorParams<float>, FactorOutputs);

template <bool SHIFT, bool FULL, typename TN>
MfmaFn<TN> pick(int n1, int m) {
#define TG_SFM(N1_, M_) \
  if (n1 == N1_ && m == M_) return factor_sumfact_mfma<N1_, M_, SHIFT, FULL, TN>;
  TG_SFM(2, 2) TG_SFM(2, 3) TG_SFM(2, 4)
  TG_SFM(3, 2) TG_SFM(3, 3) TG_SFM(3, 4)
  TG_SFM(4, 2) TG_SFM(4, 3) TG_SFM(4, 4)
#undef TG_SFM
  return nullptr;
}

// Resident workgroups per CU of the persistent launch: 2 by default (measured on Q3 96^3);
// TETRALOMA_GEOM_SF_GROUPS = 1..3 overrides it.
int groups_per_cu() {
  static const int g = [] {
    const char* e = std::getenv("TETRALOMA_GEOM_SF_GROUPS");
    const int v = e ? std::atoi(e) : 0;
    return v >= 1 && v <= 3 ? v : 2;
  }();
  return g;
}

template <typename TN>
hipError_t launch(MfmaFn<TN> fn, const TN* nodes, const uint32_t* dofmap, const TN* tab1d,
                  const uint32_t* alex, const float* tab2d, const double* tab64,
                  const FactorParams<float>& p, const FactorOutputs& out, hipStream_t stream) {
  if (!fn || p.tile_log2 != 4) return hipErrorNotSupported;
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  // groups_per_cu() workgroups per CU, no more than there are tiles.
  const uint64_t ntiles = p.padded_cells >> 4;
  uint64_t grid = static_cast<uint64_t>(cus > 0 ? cus : 8) * groups_per_cu();
  if (grid > ntiles) grid = ntiles;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(fn, dim3(static_cast<uint32_t>(grid)), dim3(64 * kWaves), 0, stream, nodes,
                     dofmap, tab1d, alex, tab2d, tab64, p, out);
  return hipGetLastError();
}

}  // namespace

bool has_sumfact_mfma(int n1, int m) { return pick<true, true, float>(n1, m) != nullptr; }

hipError_t launch_sumfact_mfma(int n1, int m, bool shift, const float* nodes,
                               const uint32_t* dofmap, const float* tab1d, const uint32_t* alex,
                               const float* tab2d, const double* tab64,
                               const FactorParams<float>& p, const FactorOutputs& out,
                               hipStream_t stream) {
  const bool full = p.mask == (kFPoints | kFJac | kFDetJ);
  MfmaFn<float> fn = shift ? (full ? pick<true, true, float>(n1, m) : pick<true, false, float>(n1, m))
                           : (full ? pick<false, true, float>(n1, m) : pick<false, false, float>(n1, m));
  return launch(fn, nodes, dofmap, tab1d, alex, tab2d, tab64, p, out, stream);
}

hipError_t launch_sumfact_mfma_f64_nodes(int n1, int m, const double* nodes,
                                         const uint32_t* dofmap, const double* tab1d,
                                         const uint32_t* alex, const float* tab2d,
                                         const double* tab64, const FactorParams<float>& p,
                                         const FactorOutputs& out, hipStream_t stream) {
  const bool full = p.mask == (kFPoints | kFJac | kFDetJ);
  MfmaFn<double> fn = full ? pick<true, true, double>(n1, m) : pick<true, false, double>(n1, m);
  return launch(fn, nodes, dofmap, tab1d, alex, tab2d, tab64, p, out, stream);
}

}  // namespace tg
