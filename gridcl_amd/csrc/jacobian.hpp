// det J, its pseudo-determinant and the pseudo-inverse K of one Jacobian J (column-major,
// J(g, d) at d * GD + g; K at g * TD + t), shared by the factor and pull-back kernels. Every
// product is rounded before it is added (no contraction into fused multiply-adds), so the values
// do not depend on the code around an inlined call: kernels of different shapes of code agree
// bit for bit.
#pragma once

#include <hip/hip_runtime.h>

namespace tg {
namespace jac {

__device__ __forceinline__ float jsqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double jsqrt(double x) { return __builtin_sqrt(x); }

// det J (tdim = gdim) or sqrt(det(J^T J)) (layout.rs:186-213).
template <typename T, int TD, int GD>
__device__ __forceinline__ T col_dot(const T* j, int a, int b) {
#pragma clang fp contract(off)
  T acc = j[a * GD] * j[b * GD];
#pragma unroll
  for (int g = 1; g < GD; ++g) acc += j[a * GD + g] * j[b * GD + g];
  return acc;
}

template <typename T, int TD, int GD>
__device__ __forceinline__ T det_of(const T* j) {
#pragma clang fp contract(off)
  if constexpr (TD == GD) {
    if constexpr (TD == 1) {
      return j[0];
    } else if constexpr (TD == 2) {
      return j[0] * j[3] - j[2] * j[1];
    } else {
      return j[0] * (j[4] * j[8] - j[7] * j[5]) - j[3] * (j[1] * j[8] - j[7] * j[2]) +
             j[6] * (j[1] * j[5] - j[4] * j[2]);
    }
  } else if constexpr (TD == 1) {
    return jsqrt(col_dot<T, TD, GD>(j, 0, 0));
  } else {
    const T a = col_dot<T, TD, GD>(j, 0, 0), b = col_dot<T, TD, GD>(j, 0, 1),
            c = col_dot<T, TD, GD>(j, 1, 1);
    return jsqrt(a * c - b * b);
  }
}

// K = (J^T J)^-1 J^T, column-major tdim x gdim (layout.rs:215-268).
template <typename T, int TD, int GD>
__device__ __forceinline__ void pinv_of(const T* j, T* k) {
#pragma clang fp contract(off)
  if constexpr (TD == GD) {
    const T inv = T(1) / det_of<T, TD, GD>(j);
    if constexpr (TD == 1) {
      k[0] = inv;
    } else if constexpr (TD == 2) {
      k[0] = j[3] * inv;
      k[2] = -j[2] * inv;
      k[1] = -j[1] * inv;
      k[3] = j[0] * inv;
    } else {
      k[0] = (j[4] * j[8] - j[7] * j[5]) * inv;
      k[3] = (j[6] * j[5] - j[3] * j[8]) * inv;
      k[6] = (j[3] * j[7] - j[6] * j[4]) * inv;
      k[1] = (j[7] * j[2] - j[1] * j[8]) * inv;
      k[4] = (j[0] * j[8] - j[6] * j[2]) * inv;
      k[7] = (j[6] * j[1] - j[0] * j[7]) * inv;
      k[2] = (j[1] * j[5] - j[4] * j[2]) * inv;
      k[5] = (j[3] * j[2] - j[0] * j[5]) * inv;
      k[8] = (j[0] * j[4] - j[3] * j[1]) * inv;
    }
  } else if constexpr (TD == 1) {
    const T inv = T(1) / col_dot<T, TD, GD>(j, 0, 0);
#pragma unroll
    for (int g = 0; g < GD; ++g) k[g] = j[g] * inv;
  } else {
    const T a = col_dot<T, TD, GD>(j, 0, 0), b = col_dot<T, TD, GD>(j, 0, 1),
            c = col_dot<T, TD, GD>(j, 1, 1);
    const T inv = T(1) / (a * c - b * b);
    const T g00 = c * inv, g01 = -b * inv, g11 = a * inv;
#pragma unroll
    for (int g = 0; g < GD; ++g) {
      k[g * TD] = g00 * j[g] + g01 * j[GD + g];
      k[g * TD + 1] = g01 * j[g] + g11 * j[GD + g];
    }
  }
}

}  // namespace jac
}  // namespace tg
