// The reference's CPU path for the geometry factors, as the benchmark's CPU baseline. TEST
// INFRASTRUCTURE ONLY (timed by bench.py's cpu_baseline leg, checked by tests/test_oracle.py);
// the library under gridcl_amd/ never links, loads or calls it.
//
// The reference runs its factor kernels on the CubeCL CPU runtime (LLVM at O3, multiply-adds
// fused: docs/adr/0013-cpu-runtime-fuses-multiply-adds-signed-zeros.md) with the CPU vector
// policy: QuadField tiles of 4 cells and GeomFactors vector size 4
// (tetraloma-device/vector-policy.toml:13-27), one launch per factor over the cached E-vector
// (tetraloma-geometry/src/layout.rs:488-772; evector.rs:197-236), and its benchmark sums the four
// timings (E-vector, X, J, det J; tetraloma-bench/src/families/degree.rs:1377-1462). This port
// keeps that schedule: the E-vector [k][3][cells padded to 4] once, then one pass per factor over
// tiles of 4 cells with the tile's cells in the four lanes of one SIMD vector, the node count
// fixed at compile time, OpenMP over tiles. det J recomputes J, as the reference's det J launch
// does. Sums start from the node-0 product and run in ascending node order (layout.rs:138-213),
// with multiply-adds fused as the reference's CPU runtime does.
#include <cmath>
#include <cstddef>
#include <cstdint>

#include <omp.h>

namespace {

typedef double v4d __attribute__((vector_size(32)));

inline v4d load4(const double* p) {
  v4d v;
  __builtin_memcpy(&v, p, sizeof v);
  return v;
}
inline void store4(double* p, v4d v) { __builtin_memcpy(p, &v, sizeof v); }

// E-vector [k][3][padded]: node k of cell c at (k * 3 + g) * padded + c; padding copies the last
// cell; indices clamped to the last node. One tile of 4 cells per iteration, so that each row
// is written as whole vectors.
void evector(int k, const double* nodes, size_t num_nodes, const uint32_t* dofmap,
             size_t num_cells, size_t padded, double* ev) {
  const long ntiles = static_cast<long>(padded / 4);
#pragma omp parallel for schedule(static)
  for (long t = 0; t < ntiles; ++t) {
    const uint32_t* rows[4];
    for (int l = 0; l < 4; ++l) {
      const size_t c = static_cast<size_t>(t) * 4 + l;
      rows[l] = dofmap + (c < num_cells ? c : num_cells - 1) * k;
    }
    for (int i = 0; i < k; ++i) {
      const double* src[4];
      for (int l = 0; l < 4; ++l) {
        uint32_t n = rows[l][i];
        if (n >= num_nodes) n = static_cast<uint32_t>(num_nodes - 1);
        src[l] = nodes + n * 3ull;
      }
      for (int g = 0; g < 3; ++g) {
        const v4d v = {src[0][g], src[1][g], src[2][g], src[3][g]};
        store4(ev + (static_cast<size_t>(i) * 3 + g) * padded + t * 4, v);
      }
    }
  }
}

// The tile's node coordinates, lanes = its 4 cells.
template <int K>
inline void tile_nodes(const double* ev, size_t padded, size_t t, v4d (&e)[K][3]) {
  for (int i = 0; i < K; ++i)
    for (int g = 0; g < 3; ++g) e[i][g] = load4(ev + (static_cast<size_t>(i) * 3 + g) * padded + t * 4);
}

// X = sum_k x_k phi_k + origin, [tile][q][3][4].
template <int K>
void x_pass(const double* ev, size_t padded, const double* phi, size_t nq, const double* origin,
            double* x) {
  const long ntiles = static_cast<long>(padded / 4);
#pragma omp parallel for schedule(static)
  for (long t = 0; t < ntiles; ++t) {
    v4d e[K][3];
    tile_nodes<K>(ev, padded, t, e);
    for (size_t q = 0; q < nq; ++q) {
      const double* f = phi + q * K;
      for (int g = 0; g < 3; ++g) {
        v4d acc = e[0][g] * f[0];
        for (int i = 1; i < K; ++i) acc += e[i][g] * f[i];
        store4(x + ((t * nq + q) * 3 + g) * 4, acc + origin[g]);
      }
    }
  }
}

template <int K>
inline void jacobian(const v4d (&e)[K][3], const double* d, v4d (&jm)[9]) {
  for (int dd = 0; dd < 3; ++dd)
    for (int g = 0; g < 3; ++g) {
      v4d acc = e[0][g] * d[dd];
      for (int i = 1; i < K; ++i) acc += e[i][g] * d[i * 3 + dd];
      jm[dd * 3 + g] = acc;
    }
}

// J (column-major, component d * 3 + g), [tile][q][9][4].
template <int K>
void j_pass(const double* ev, size_t padded, const double* dphi, size_t nq, double* j) {
  const long ntiles = static_cast<long>(padded / 4);
#pragma omp parallel for schedule(static)
  for (long t = 0; t < ntiles; ++t) {
    v4d e[K][3];
    tile_nodes<K>(ev, padded, t, e);
    for (size_t q = 0; q < nq; ++q) {
      v4d jm[9];
      jacobian<K>(e, dphi + q * K * 3, jm);
      for (int i = 0; i < 9; ++i) store4(j + ((t * nq + q) * 9 + i) * 4, jm[i]);
    }
  }
}

// det J by first-row cofactors of the column-major J (layout.rs:186-213), [tile][q][4].
template <int K>
void detj_pass(const double* ev, size_t padded, const double* dphi, size_t nq, double* detj) {
  const long ntiles = static_cast<long>(padded / 4);
#pragma omp parallel for schedule(static)
  for (long t = 0; t < ntiles; ++t) {
    v4d e[K][3];
    tile_nodes<K>(ev, padded, t, e);
    for (size_t q = 0; q < nq; ++q) {
      v4d a[9];
      jacobian<K>(e, dphi + q * K * 3, a);
      const v4d det = a[0] * (a[4] * a[8] - a[7] * a[5]) - a[3] * (a[1] * a[8] - a[7] * a[2]) +
                      a[6] * (a[1] * a[5] - a[4] * a[2]);
      store4(detj + (t * nq + q) * 4, det);
    }
  }
}

template <int K>
void factors(unsigned passes, const double* ev, size_t padded, const double* phi,
             const double* dphi, size_t nq, const double* origin, double* x, double* j,
             double* detj) {
  if (passes & 2u) x_pass<K>(ev, padded, phi, nq, origin, x);
  if (passes & 4u) j_pass<K>(ev, padded, dphi, nq, j);
  if (passes & 8u) detj_pass<K>(ev, padded, dphi, nq, detj);
}

}  // namespace

extern "C" {

// Tetrahedra in 3D with k = 4, 10 or 20 nodes (P1-P3): nodes [n][3] relative to origin, dofmap
// [c][k], tables phi [nq][k] and dphi [nq][k][3]; evector of 3 k padded doubles, x [padded nq 3],
// j [padded nq 9], detj [padded nq], padded = cells rounded up to 4. `passes` selects the
// launches, so that each can be timed alone as the reference's benchmark does: bit 0 the
// E-vector, bit 1 X, bit 2 J, bit 3 det J (the last three read the E-vector left by an earlier
// call). threads > 0 sets OpenMP's thread count. Returns 0, or -1 for another k.
int cpu_port_factors(int k, unsigned passes, const double* nodes, size_t num_nodes,
                     const double* origin, const uint32_t* dofmap, size_t num_cells,
                     const double* phi, const double* dphi, size_t nq, double* ev, double* x,
                     double* j, double* detj, int threads) {
  if (k != 4 && k != 10 && k != 20) return -1;
  if (threads > 0) omp_set_num_threads(threads);
  const size_t padded = (num_cells + 3) / 4 * 4;
  if (passes & 1u) evector(k, nodes, num_nodes, dofmap, num_cells, padded, ev);
  if (k == 4) factors<4>(passes, ev, padded, phi, dphi, nq, origin, x, j, detj);
  if (k == 10) factors<10>(passes, ev, padded, phi, dphi, nq, origin, x, j, detj);
  if (k == 20) factors<20>(passes, ev, padded, phi, dphi, nq, origin, x, j, detj);
  return 0;
}

}  // extern "C"
