// The coordinate map of one cell at a point xi that changes at run time, for the pull-back:
// X_g(xi) = sum_r A(r, g) f_r(xi) and J(g, d) = sum_r A(r, g) d f_r / d xi_d, where f_r is the
// spanning set of the cell's Lagrange space and A = C^T x the cell's coefficients in it (C the
// inverse Vandermonde on the element's nodes, computed on the host).
//
// The spanning sets are those of the host element library (refcell.cpp), evaluated by the same
// recurrences: Legendre products L_a(2x - 1) L_b(2y - 1) L_c(2z - 1) on intervals,
// quadrilaterals and hexahedra; the collapsed (Dubiner) bases on triangles and tetrahedra,
// S_i(2x + y + z - 1, 1 - y - z) S^(2i+1)_j(2y + z - 1, 1 - z) P^(2i+2j+2)_k(2z - 1) with
// S^(a)_n(x, s) = s^n P^(a,0)_n(x / s); the triangle basis times P_k(2z - 1) on prisms; the
// rational monomials x^a y^b z^c / (1 - z)^min(a,b) on pyramids (tetraloma-geometry/src/
// basis_eval/mod.rs:1-24 evaluates the reference's orthogonal sets the same way).
//
// The order r of the functions is the loop order below, which the host reproduces when it lays
// out C (pullback.cpp, device_order): tensor cells c, b, a (a fastest); triangles i, j;
// tetrahedra i, j, k; prisms i, j, k (k fastest); pyramids c, b, a. Every loop has compile-time
// bounds, so r and every array index are constants after unrolling.
#pragma once

#include <hip/hip_runtime.h>

#include "refcell.hpp"

namespace tg {
namespace span {

// Legendre L_n(2t - 1), n = 0..P, and their t-derivatives (refcell.cpp, legendre01).
template <typename T, int P>
__device__ __forceinline__ void legendre01(T t, T (&v)[P + 1], T (&dv)[P + 1]) {
  const T x = T(2) * t - T(1);
  v[0] = T(1);
  dv[0] = T(0);
  if constexpr (P >= 1) {
    v[1] = x;
    dv[1] = T(2);
  }
#pragma unroll
  for (int k = 2; k <= P; ++k) {
    const T a = T(2 * k - 1), b = T(k - 1), inv = T(1) / T(k);
    v[k] = (a * x * v[k - 1] - b * v[k - 2]) * inv;
    dv[k] = (a * (T(2) * v[k - 1] + x * dv[k - 1]) - b * dv[k - 2]) * inv;
  }
}

// Scaled Jacobi S_n(x, s) = s^n P^(a,0)_n(x / s), n = 0..n_max (n_max <= P), with the partial
// derivatives in x and s (refcell.cpp, scaled_jacobi).
template <typename T, int P>
__device__ __forceinline__ void scaled_jacobi(int n_max, T a, T x, T s, T (&v)[P + 1],
                                              T (&dx)[P + 1], T (&ds)[P + 1]) {
  v[0] = T(1);
  dx[0] = T(0);
  ds[0] = T(0);
  if (n_max >= 1) {
    v[1] = T(0.5) * (a * s + (a + T(2)) * x);
    dx[1] = T(0.5) * (a + T(2));
    ds[1] = T(0.5) * a;
  }
#pragma unroll
  for (int n = 2; n <= P; ++n) {
    if (n > n_max) break;
    const T c = T(2 * n) + a;
    const T A = (c - T(1)) * c * (c - T(2));
    const T B = (c - T(1)) * a * a;
    const T D = T(2) * (T(n) + a - T(1)) * T(n - 1) * c;
    const T inv = T(1) / (T(2 * n) * (T(n) + a) * (c - T(2)));
    const T l = A * x + B * s;
    v[n] = (l * v[n - 1] - D * s * s * v[n - 2]) * inv;
    dx[n] = (A * v[n - 1] + l * dx[n - 1] - D * s * s * dx[n - 2]) * inv;
    ds[n] = (B * v[n - 1] + l * ds[n - 1] - T(2) * D * s * v[n - 2] - D * s * s * ds[n - 2]) * inv;
  }
}

// A(r, g) of the cell at coefficient row r: a[(r * GD + g) * 16].
template <typename T, int GD>
__device__ __forceinline__ T coef_at(const T* __restrict__ a, int r, int g) {
  return a[(r * GD + g) * 16];
}

// X(xi) and J(xi) of one cell of type CELL and degree P. The sums run innermost index first
// (partial sums over the fastest spanning index, then the next), so each term costs two
// multiply-adds per coordinate.
template <typename T, int CELL, int TD, int GD, int P>
__device__ __forceinline__ void coordinate_map(const T* __restrict__ a, const T (&xi)[TD],
                                               T (&X)[GD], T (&J)[TD * GD]) {
#pragma unroll
  for (int g = 0; g < GD; ++g) X[g] = T(0);
#pragma unroll
  for (int i = 0; i < TD * GD; ++i) J[i] = T(0);
  if constexpr (CELL == kInterval || CELL == kQuadrilateral || CELL == kHexahedron) {
    T v[3][P + 1], dv[3][P + 1];
#pragma unroll
    for (int d = 0; d < TD; ++d) legendre01<T, P>(xi[d], v[d], dv[d]);
    int r = 0;
#pragma unroll
    for (int c = 0; c <= (TD >= 3 ? P : 0); ++c) {
      T Sc[GD], Scx[GD], Scy[GD];
#pragma unroll
      for (int g = 0; g < GD; ++g) Sc[g] = Scx[g] = Scy[g] = T(0);
#pragma unroll
      for (int b = 0; b <= (TD >= 2 ? P : 0); ++b) {
        T Sb[GD], Sbx[GD];
#pragma unroll
        for (int g = 0; g < GD; ++g) Sb[g] = Sbx[g] = T(0);
#pragma unroll
        for (int aa = 0; aa <= P; ++aa, ++r)
#pragma unroll
          for (int g = 0; g < GD; ++g) {
            const T cf = coef_at<T, GD>(a, r, g);
            Sb[g] += cf * v[0][aa];
            Sbx[g] += cf * dv[0][aa];
          }
#pragma unroll
        for (int g = 0; g < GD; ++g) {
          if constexpr (TD == 1) {
            X[g] = Sb[g];
            J[g] = Sbx[g];
          } else {
            Sc[g] += Sb[g] * v[1][b];
            Scx[g] += Sbx[g] * v[1][b];
            Scy[g] += Sb[g] * dv[1][b];
          }
        }
      }
      if constexpr (TD == 2) {
#pragma unroll
        for (int g = 0; g < GD; ++g) {
          X[g] = Sc[g];
          J[g] = Scx[g];
          J[GD + g] = Scy[g];
        }
      } else if constexpr (TD == 3) {
#pragma unroll
        for (int g = 0; g < GD; ++g) {
          X[g] += Sc[g] * v[2][c];
          J[g] += Scx[g] * v[2][c];
          J[GD + g] += Scy[g] * v[2][c];
          J[2 * GD + g] += Sc[g] * dv[2][c];
        }
      }
    }
  } else if constexpr (CELL == kTriangle || CELL == kTetrahedron || CELL == kPrism) {
    // f = A_i B_j C_k with A = S_i(u, s), B = S^(2i+1)_j, C = S^(2i+2j+2)_k (tetrahedron) or
    // P_k(2z - 1) (prism) or 1 (triangle); with s0 = sum_k A(r) C_k and s1 = sum_k A(r) dC_k/dz:
    //   X += A B s0, dX/dx += 2 A_u B s0, dX/dy += ((A_u - A_s) B + 2 A B_u) s0,
    //   dX/dz += ((A_u - A_s) B + A (B_u - B_s)) s0 + A B s1 (tetrahedron), A B s1 (prism).
    constexpr bool tet = CELL == kTetrahedron;
    const T x = xi[0], y = xi[1], z = TD == 3 ? xi[TD - 1] : T(0);
    T A[P + 1], Au[P + 1], As[P + 1];
    scaled_jacobi<T, P>(P, T(0), tet ? T(2) * x + y + z - T(1) : T(2) * x + y - T(1),
                        tet ? T(1) - y - z : T(1) - y, A, Au, As);
    T L[P + 1], Lx[P + 1], Ls[P + 1];  // the prism's P_k(2z - 1)
    if constexpr (CELL == kPrism) scaled_jacobi<T, P>(P, T(0), T(2) * z - T(1), T(1), L, Lx, Ls);
    int r = 0;
#pragma unroll
    for (int i = 0; i <= P; ++i) {
      T B[P + 1], Bu[P + 1], Bs[P + 1];
      if constexpr (tet)
        scaled_jacobi<T, P>(P - i, T(2 * i + 1), T(2) * y + z - T(1), T(1) - z, B, Bu, Bs);
      else
        scaled_jacobi<T, P>(P - i, T(2 * i + 1), T(2) * y - T(1), T(1), B, Bu, Bs);
#pragma unroll
      for (int j = 0; j <= P - i; ++j) {
        T s0[GD], s1[GD];
#pragma unroll
        for (int g = 0; g < GD; ++g) s0[g] = s1[g] = T(0);
        if constexpr (tet) {
          T C[P + 1], Cu[P + 1], Cs[P + 1];
          scaled_jacobi<T, P>(P - i - j, T(2 * i + 2 * j + 2), T(2) * z - T(1), T(1), C, Cu, Cs);
#pragma unroll
          for (int k = 0; k <= P - i - j; ++k, ++r)
#pragma unroll
            for (int g = 0; g < GD; ++g) {
              const T cf = coef_at<T, GD>(a, r, g);
              s0[g] += cf * C[k];
              s1[g] += cf * (T(2) * Cu[k]);
            }
        } else if constexpr (CELL == kPrism) {
#pragma unroll
          for (int k = 0; k <= P; ++k, ++r)
#pragma unroll
            for (int g = 0; g < GD; ++g) {
              const T cf = coef_at<T, GD>(a, r, g);
              s0[g] += cf * L[k];
              s1[g] += cf * (T(2) * Lx[k]);
            }
        } else {
#pragma unroll
          for (int g = 0; g < GD; ++g) s0[g] = coef_at<T, GD>(a, r, g);
          ++r;
        }
        const T ab = A[i] * B[j];
        const T fx = T(2) * Au[i] * B[j];
        const T dA = (Au[i] - As[i]) * B[j];
        const T fy = dA + A[i] * (T(2) * Bu[j]);
#pragma unroll
        for (int g = 0; g < GD; ++g) {
          X[g] += ab * s0[g];
          J[g] += fx * s0[g];
          J[GD + g] += fy * s0[g];
          if constexpr (tet) J[2 * GD + g] += (dA + A[i] * (Bu[j] - Bs[j])) * s0[g] + ab * s1[g];
          if constexpr (CELL == kPrism) J[2 * GD + g] += ab * s1[g];
        }
      }
    }
  } else {  // pyramid
    const T x = xi[0], y = xi[1], z = xi[2], s = T(1) - z;
    T px[P + 1], py[P + 1], pz[P + 1];
    px[0] = py[0] = pz[0] = T(1);
#pragma unroll
    for (int e = 1; e <= P; ++e) {
      px[e] = px[e - 1] * x;
      py[e] = py[e - 1] * y;
      pz[e] = pz[e - 1] * z;
    }
    int r = 0;
#pragma unroll
    for (int c = 0; c <= P; ++c)
#pragma unroll
      for (int b = 0; b <= P - c; ++b)
#pragma unroll
        for (int aa = 0; aa <= P - c; ++aa, ++r) {
          const int n = aa < b ? aa : b;
          T rr = T(1), rz = T(0);
          bool zero = false;
          if (n > 0) {
            if (s == T(0)) {
              zero = true;
            } else {
              T sn = s;
#pragma unroll
              for (int e = 1; e < P; ++e)
                if (e < n) sn *= s;
              rr = T(1) / sn;
              rz = T(n) * rr / s;
            }
          }
          const T dxa = aa ? T(aa) * px[aa - 1] : T(0);
          const T dyb = b ? T(b) * py[b - 1] : T(0);
          const T dzc = c ? T(c) * pz[c - 1] : T(0);
          const T f = zero ? T(0) : px[aa] * py[b] * pz[c] * rr;
          const T fx = zero ? T(0) : dxa * py[b] * pz[c] * rr;
          const T fy = zero ? T(0) : px[aa] * dyb * pz[c] * rr;
          const T fz = zero ? T(0) : px[aa] * py[b] * (dzc * rr + pz[c] * rz);
#pragma unroll
          for (int g = 0; g < GD; ++g) {
            const T cf = coef_at<T, GD>(a, r, g);
            X[g] += cf * f;
            J[g] += cf * fx;
            J[GD + g] += cf * fy;
            J[2 * GD + g] += cf * fz;
          }
        }
  }
}

}  // namespace span
}  // namespace tg
