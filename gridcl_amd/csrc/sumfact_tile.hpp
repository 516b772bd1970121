// The sum-factorised evaluation of one 16-cell tile of hexahedra by one 256-thread workgroup, as a
// device function over caller-provided LDS (the per-block kernel in sumfact_kernels.hip calls
// it). The arithmetic is described in sumfact_kernels.hip, the schedule at tile() below.
#pragma once

#include <hip/hip_runtime.h>

#include "geom_internal.hpp"

namespace tg {
namespace sf {

constexpr int kThreads = 256;  // four wavefronts per tile of 16 cells

__device__ __forceinline__ float tfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double tfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float tsqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double tsqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float tabs(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ double tabs(double x) { return __builtin_fabs(x); }

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

template <typename T>
struct Vec2 {
  typedef T type __attribute__((ext_vector_type(2)));
};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// FP64 det J of cell c at point q from the stored coordinates (of type TN; offsets from node 0
// in FP64), for near-degenerate points of FP32 geometries.
template <int K, typename TN>
__device__ __noinline__ double detj_fp64(const TN* __restrict__ nodes,
                                         const uint32_t* __restrict__ row, uint32_t last_node,
                                         const double* __restrict__ tab64, uint32_t q) {
  double x0[3], j[9];
  const uint32_t n0 = row[0] < last_node ? row[0] : last_node;
#pragma unroll
  for (int g = 0; g < 3; ++g) x0[g] = static_cast<double>(nodes[static_cast<uint64_t>(n0) * 3 + g]);
#pragma unroll
  for (int i = 0; i < 9; ++i) j[i] = 0.0;
  for (int k = 1; k < K; ++k) {
    const uint32_t n = row[k] < last_node ? row[k] : last_node;
    const double* t = tab64 + (static_cast<uint64_t>(q) * K + k) * 4;
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

// Copies n elements (a multiple of 16 bytes) from LDS to global memory with 16-byte
// non-temporal stores spread over the wavefront's lanes.
template <typename T>
__device__ __forceinline__ void flush_run(const T* __restrict__ src, T* __restrict__ dst, int n,
                                          uint32_t lane) {
  using V = typename Vec16<T>::type;
  constexpr int E = 16 / sizeof(T);
  for (int i = static_cast<int>(lane); i < n / E; i += 64) {
    const V v = *reinterpret_cast<const V*>(src + i * E);
    __builtin_nontemporal_store(v, reinterpret_cast<V*>(dst + i * E));
  }
}

// Lane v's value of x (a wavefront-uniform copy).
__device__ __forceinline__ uint32_t readlane_bits(float x, int v) {
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(__builtin_bit_cast(int, x), v));
}
__device__ __forceinline__ uint64_t readlane_bits(double x, int v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(b), v));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(b >> 32), v));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// LDS of one tile (elements of T): the gathered nodes [3][16][K + 1] (reused by the output
// stages, one per wavefront, once the first contraction has read them), the cells' vertex means
// [3][16], and the first contraction's sums [3 g][N1 a][N1 b][M r][16 cells][2] (rows of one
// (g, a) padded by 32 elements, so that the four a of one wavefront's writes fall in two bank
// halves).
template <int N1, int M, int PC>
struct Lds {
  static constexpr int K = N1 * N1 * N1;
  static constexpr int kXs = 3 * 16 * (K + 1);
  static constexpr int kStage = PC * 13 * 16;  // per wavefront
  static constexpr int kFront = kXs > 4 * kStage ? kXs : 4 * kStage;
  static constexpr int kShift = 3 * 16;
  static constexpr int kRowA = N1 * M * 32 + 32;
  static constexpr int kSums = 3 * N1 * kRowA;
  static constexpr int kTotal = kFront + kShift + kSums;
  static_assert(kFront % 4 == 0 && kStage % 4 == 0, "16-byte aligned stages and sums");
};

// One tile: N1 nodes and M points per direction; PC points of the p direction per pass (PC < M
// recomputes the first two sums to save registers, for FP64).
//
// One workgroup of four wavefronts per tile of 16 cells. Gather: each wavefront loads four of the
// tile's cells, one node per lane, through the dofmap (indices clamped to the last node) and,
// for FP32, subtracts each cell's vertex mean (read from lanes 0-7, no barrier) before
// staging them in LDS as [g][16 cells][lex node] with rows of K + 1 (the gather's writes and the
// first contraction's reads both fall in distinct banks).
// The first contraction (along zeta) is shared by the workgroup: its sums for every (cell, g,
// a, b, r) go to LDS once, instead of once per wavefront.
// Wavefront w takes the points with q = w; its lanes are 16 cells x 4 values of r, and each lane
// contracts those sums along eta and xi into X, J and det J at its points (p, w, r), for all
// three coordinates at once. The results leave through a per-wavefront LDS stage: the points (0..M-1, q, r) are consecutive in
// the point order, so X, J and det J of one r are three contiguous runs of the QuadField (tiles
// of 16), written with 16-byte stores in order; X and J go one r at a time, then det J (after
// any FP64 recomputation) of all r.
// `lds` holds Lds<N1, M, PC>::kTotal elements (16-byte aligned); every thread of the 256-thread
// workgroup calls with the same tile.
// TN is the type the node coordinates are stored in: T, or double for FP32 arithmetic on FP64
// nodes (with SHIFT; the vertex mean and the offsets are then taken in FP64 and rounded once).
// TO is the type of the outputs and of the launch's parameters: T, or float for FP64 arithmetic
// with FP32 outputs (each value rounded once, at its store).
template <typename T, int N1, int M, int PC, bool SHIFT, typename TN = T, typename TO = T>
__device__ __forceinline__ void tile(const TN* __restrict__ nodes, const uint32_t* __restrict__ dofmap,
                                     const T* __restrict__ tab1d, const uint32_t* __restrict__ lex,
                                     const double* __restrict__ tab64, const FactorParams<TO>& p,
                                     const FactorOutputs& o, uint64_t tile, T* __restrict__ lds) {
  using L = Lds<N1, M, PC>;
  constexpr int K = N1 * N1 * N1;
  constexpr int NQ = M * M * M;
  constexpr int kBlock = kThreads;
  constexpr int SX = 0, SJ = PC * 3 * 16, SLEN = L::kStage;
  static_assert(M * PC * 16 <= SLEN, "det J of every r fits one stage");
  using T2 = typename Vec2<T>::type;
  T* __restrict__ xs = lds;
  T* __restrict__ stage_base = lds;  // after the first contraction
  T* __restrict__ shift = lds + L::kFront;
  T* __restrict__ sums = lds + L::kFront + L::kShift;
  const uint32_t t = threadIdx.x, lane = t & 63u, cell = t & 15u, r_lane = (t >> 4) & 3u,
                 w = t >> 6;

  static_assert(sizeof(TN) == sizeof(T) || SHIFT, "wider node storage needs the shift");
  struct P3 {
    TN v[3];
  };
  // Gather: wavefront w takes the cells w, w + 4, w + 8, w + 12 of the tile, lane k < K node k
  // (A.5 order: the eight vertices are lanes 0-7).
  TN held[4][3];
  const bool node_lane = lane < static_cast<uint32_t>(K);
  const uint32_t lex_lane = node_lane ? lex[lane] : 0u;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t cc = w + 4u * i;
    uint64_t c = tile * 16 + cc;
    c = c < p.num_cells ? c : p.num_cells - 1;
    held[i][0] = held[i][1] = held[i][2] = TN(0);
    if (node_lane) {
      uint32_t n = dofmap[c * K + lane];
      n = n < p.last_node ? n : p.last_node;
      const P3 x = *reinterpret_cast<const P3*>(nodes + static_cast<uint64_t>(n) * 3);
#pragma unroll
      for (int g = 0; g < 3; ++g) held[i][g] = x.v[g];
    }
  }
  // FP32: each cell's vertex mean, summed in vertex order from lanes 0-7 (wavefront-uniform),
  // subtracted before the nodes are staged. (A pairwise DPP sum, or 12 lanes summing from LDS,
  // issue fewer instructions but measured slower.)
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t cc = w + 4u * i;
#pragma unroll
    for (int g = 0; g < 3; ++g) {
      TN sg = TN(0);
      if constexpr (SHIFT) {
        TN acc = __builtin_bit_cast(TN, readlane_bits(held[i][g], 0));
#pragma unroll
        for (int v = 1; v < 8; ++v) acc += __builtin_bit_cast(TN, readlane_bits(held[i][g], v));
        sg = acc * TN(0.125);
        if (lane == 0) shift[g * 16 + cc] = static_cast<T>(sg);
      }
      if (node_lane) xs[(g * 16 + cc) * (K + 1) + lex_lane] = static_cast<T>(held[i][g] - sg);
    }
  }
  __syncthreads();
  // First contraction (along zeta, the node index c): for each (cell, g, a, b) and each r,
  // s0 = sum_c x b_r(c) and s1 = sum_c x b_r'(c). Items are spread over the workgroup with the
  // cell fastest; for N1 = 4 the four 16-lane groups of a wavefront take the four a of one b,
  // so that their reads of the node rows fall in distinct banks.
  {
    constexpr int kItems = 16 * 3 * N1 * N1;
    constexpr int kPer = (kItems + kBlock - 1) / kBlock;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const uint32_t it = t + kBlock * i;
      if (kItems % kBlock == 0 || it < static_cast<uint32_t>(kItems)) {
        const uint32_t cc = it & 15u, rest = it >> 4;
        const uint32_t g = rest / (N1 * N1), u = rest - g * (N1 * N1), a = u % N1, b = u / N1;
        const T* __restrict__ xr = xs + (g * 16 + cc) * (K + 1) + (a * N1 + b) * N1;
        T xv[N1];
#pragma unroll
        for (int c = 0; c < N1; ++c) xv[c] = xr[c];
        T* __restrict__ dst = sums + (g * N1 + a) * L::kRowA + b * (M * 32) + cc * 2;
#pragma unroll
        for (int rr = 0; rr < M; ++rr) {
          T s0 = T(0), s1 = T(0);
#pragma unroll
          for (int c = 0; c < N1; ++c) {
            s0 = tfma(xv[c], tab1d[rr * N1 + c], s0);
            s1 = tfma(xv[c], tab1d[(M + rr) * N1 + c], s1);
          }
          *reinterpret_cast<T2*>(dst + rr * 32) = T2{s0, s1};
        }
      }
    }
  }
  __syncthreads();
  const int q = static_cast<int>(w);
  if (q >= M) return;  // the whole wavefront; no barrier follows in this function
  // Lanes with r >= M compute r = M - 1 and stage nothing.
  const bool r_live = r_lane < static_cast<uint32_t>(M);
  const uint32_t r = r_live ? r_lane : M - 1;
  // The output stage holds TO values (no wider than T) in the same LDS.
  TO* __restrict__ stage = reinterpret_cast<TO*>(stage_base + w * SLEN);
  T sg[3];
#pragma unroll
  for (int g = 0; g < 3; ++g) sg[g] = SHIFT ? shift[g * 16 + cell] : T(0);
  const T* __restrict__ s_lane = sums + (r * 16 + cell) * 2;
  const uint32_t mask = p.mask;
  uint64_t c_own = tile * 16 + cell;
  c_own = c_own < p.num_cells ? c_own : p.num_cells - 1;

#pragma unroll
  for (int p0 = 0; p0 < M; p0 += PC) {
    constexpr int kPC = PC;
    const int npp = M - p0 < kPC ? M - p0 : kPC;
    // acc[pp][v][g]: v = 0 X, 1..3 the derivatives along xi, eta, zeta. The sums over b for one
    // a (s2[w][g]: w = 0 value-value, 1 d eta, 2 d zeta), then their terms of the sums over a;
    // a = 0 starts the sums (a product, as a fused multiply-add onto zero would give).
    T acc[PC][4][3];
    auto sums_b = [&](int a, T (&s2)[3][3]) {
#pragma unroll
      for (int i = 0; i < 3; ++i) s2[i][0] = s2[i][1] = s2[i][2] = T(0);
#pragma unroll
      for (int b = 0; b < N1; ++b) {
        const T bq = tab1d[q * N1 + b], dq = tab1d[(M + q) * N1 + b];
#pragma unroll
        for (int g = 0; g < 3; ++g) {
          const T2 sv = *reinterpret_cast<const T2*>(s_lane + (g * N1 + a) * L::kRowA + b * (M * 32));
          s2[0][g] = tfma(sv.x, bq, s2[0][g]);
          s2[1][g] = tfma(sv.x, dq, s2[1][g]);
          s2[2][g] = tfma(sv.y, bq, s2[2][g]);
        }
      }
    };
#pragma unroll
    for (int pp = 0; pp < PC; ++pp)
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int g = 0; g < 3; ++g) acc[pp][v][g] = T(0);
#pragma unroll 1
    for (int a = 0; a < N1; ++a) {
      T s2[3][3];
      sums_b(a, s2);
#pragma unroll
      for (int pp = 0; pp < PC; ++pp) {
        if (p0 + pp >= M) continue;
        const T ba = tab1d[(p0 + pp) * N1 + a], da = tab1d[(M + p0 + pp) * N1 + a];
#pragma unroll
        for (int g = 0; g < 3; ++g) {
          acc[pp][0][g] = tfma(s2[0][g], ba, acc[pp][0][g]);
          acc[pp][1][g] = tfma(s2[0][g], da, acc[pp][1][g]);
          acc[pp][2][g] = tfma(s2[1][g], ba, acc[pp][2][g]);
          acc[pp][3][g] = tfma(s2[2][g], ba, acc[pp][3][g]);
        }
      }
    }
    // det J (first-row cofactor expansion of the column-major J, layout.rs:186-213). FP32
    // points with |det J| < tau |J|_F^3 are marked in `near` and recomputed in FP64 once X and J
    // have left (so that the accumulators are not live across that call).
    T det[PC];
    uint32_t near = 0;
#pragma unroll
    for (int pp = 0; pp < PC; ++pp) {
      const T(&A)[4][3] = acc[pp];
      // J(g, d) = A[1 + d][g]
      det[pp] = A[1][0] * (A[2][1] * A[3][2] - A[3][1] * A[2][2]) -
                A[2][0] * (A[1][1] * A[3][2] - A[3][1] * A[1][2]) +
                A[3][0] * (A[1][1] * A[2][2] - A[2][1] * A[1][2]);
      if constexpr (sizeof(T) == 4) {
        if (p.fixup && p0 + pp < M) {
          T n2 = T(0);
#pragma unroll
          for (int d = 1; d < 4; ++d)
#pragma unroll
            for (int g = 0; g < 3; ++g) n2 = tfma(A[d][g], A[d][g], n2);
          if (tabs(det[pp]) < static_cast<T>(p.tau) * n2 * tsqrt(n2)) near |= 1u << pp;
        }
      }
    }
    // X and J, one r at a time: stage the points (p0.., q, r) and write their two runs.
    if (mask & (kFPoints | kFJac)) {
#pragma unroll 1
      for (int rr = 0; rr < M; ++rr) {
        wave_sync();
        if (r_live && r == static_cast<uint32_t>(rr)) {
#pragma unroll
          for (int pp = 0; pp < PC; ++pp) {
            if (p0 + pp >= M) continue;
#pragma unroll
            for (int g = 0; g < 3; ++g) {
              stage[SX + (pp * 3 + g) * 16 + cell] =
                  static_cast<TO>((acc[pp][0][g] + sg[g]) + static_cast<T>(p.origin[g]));
#pragma unroll
              for (int d = 0; d < 3; ++d)
                stage[SJ + (pp * 9 + d * 3 + g) * 16 + cell] = static_cast<TO>(acc[pp][1 + d][g]);
            }
          }
        }
        wave_sync();
        const uint64_t pt0 = tile * NQ + static_cast<uint32_t>(p0 + M * q + M * M * rr);
        if (mask & kFPoints)
          flush_run<TO>(stage + SX, static_cast<TO*>(o.x) + ((pt0 * 3) << 4), npp * 3 * 16, lane);
        if (mask & kFJac)
          flush_run<TO>(stage + SJ, static_cast<TO*>(o.j) + ((pt0 * 9) << 4), npp * 9 * 16, lane);
      }
    }
    if (mask & kFDetJ) {
      if constexpr (sizeof(T) == 4) {
        if (near) {
#pragma unroll
          for (int pp = 0; pp < PC; ++pp)
            if (near & (1u << pp))
              det[pp] = static_cast<T>(detj_fp64<K>(
                  nodes, dofmap + c_own * K, p.last_node, tab64,
                  static_cast<uint32_t>(p0 + pp + M * q) + M * M * r));
        }
      }
      // det J of all r at once: run rr at stage + rr * PC * 16.
      wave_sync();
      if (r_live)
#pragma unroll
        for (int pp = 0; pp < PC; ++pp)
          if (p0 + pp < M) stage[(r * PC + pp) * 16 + cell] = static_cast<TO>(det[pp]);
      wave_sync();
#pragma unroll 1
      for (int rr = 0; rr < M; ++rr) {
        const uint64_t pt0 = tile * NQ + static_cast<uint32_t>(p0 + M * q + M * M * rr);
        flush_run<TO>(stage + rr * PC * 16, static_cast<TO*>(o.detj) + (pt0 << 4), npp * 16, lane);
      }
    }
  }
}

}  // namespace sf
}  // namespace tg
