// C-ABI entry point of the pull-back (tetraloma-geometry/src/pull_back.rs:566, pull_back_block).
#include <hip/hip_runtime.h>

#include <array>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tetraloma_geom.h"
#include "handles.hpp"
#include "pullback_internal.hpp"

namespace {

using tg::Cell;

// The spanning functions (a, b, c) in the order the kernels' loops visit them
// (spanning.hpp, coordinate_map).
std::vector<std::array<int, 3>> device_order(Cell cell, int p) {
  std::vector<std::array<int, 3>> out;
  switch (cell) {
    case tg::kInterval:
      for (int a = 0; a <= p; ++a) out.push_back({a, 0, 0});
      break;
    case tg::kQuadrilateral:
      for (int b = 0; b <= p; ++b)
        for (int a = 0; a <= p; ++a) out.push_back({a, b, 0});
      break;
    case tg::kHexahedron:
      for (int c = 0; c <= p; ++c)
        for (int b = 0; b <= p; ++b)
          for (int a = 0; a <= p; ++a) out.push_back({a, b, c});
      break;
    case tg::kTriangle:
      for (int i = 0; i <= p; ++i)
        for (int j = 0; j <= p - i; ++j) out.push_back({i, j, 0});
      break;
    case tg::kTetrahedron:
      for (int i = 0; i <= p; ++i)
        for (int j = 0; j <= p - i; ++j)
          for (int k = 0; k <= p - i - j; ++k) out.push_back({i, j, k});
      break;
    case tg::kPrism:
      for (int i = 0; i <= p; ++i)
        for (int j = 0; j <= p - i; ++j)
          for (int k = 0; k <= p; ++k) out.push_back({i, j, k});
      break;
    case tg::kPyramid:
      for (int c = 0; c <= p; ++c)
        for (int b = 0; b <= p - c; ++b)
          for (int a = 0; a <= p - c; ++a) out.push_back({a, b, c});
      break;
    default:
      break;
  }
  return out;
}

// The element's C [r][k] (row r: the coefficients of spanning function r in the Lagrange basis
// functions k) in the kernels' order, narrowed to the geometry's type and uploaded once per
// geometry and element.
int coefficients(tg_geometry* g, const tg::Block& b, const void** out) {
  const auto key = std::make_tuple(static_cast<int>(b.cell), b.degree, static_cast<int>(b.family));
  auto hit = g->pull_back_coef.find(key);
  if (hit != g->pull_back_coef.end()) {
    *out = hit->second;
    return TG_OK;
  }
  tg::LagrangeElement el;
  std::string err;
  if (!el.init(b.cell, b.degree, b.family, &err))
    return tg::fail(TG_ERR_UNSUPPORTED_NODE_FAMILY, "element: " + err);
  const size_t dim = el.dim();
  const std::vector<int>& exps = el.exps();
  const std::vector<std::array<int, 3>> order = device_order(b.cell, static_cast<int>(b.degree));
  if (order.size() != dim) return tg::fail(TG_ERR_UNSUPPORTED_DEGREE, "UnsupportedDegree");
  std::vector<double> c(dim * dim);
  for (size_t r = 0; r < dim; ++r) {
    size_t jj = dim;
    for (size_t j = 0; j < dim; ++j)
      if (exps[3 * j] == order[r][0] && exps[3 * j + 1] == order[r][1] &&
          exps[3 * j + 2] == order[r][2])
        jj = j;
    if (jj == dim) return tg::fail(TG_ERR_UNSUPPORTED_DEGREE, "internal: spanning order");
    for (size_t k = 0; k < dim; ++k) c[r * dim + k] = el.coef()[jj * dim + k];
  }
  const size_t elem = g->dtype == TG_F32 ? sizeof(float) : sizeof(double);
  std::vector<char> bytes(dim * dim * elem);
  for (size_t i = 0; i < dim * dim; ++i) {
    if (g->dtype == TG_F32) {
      const float v = static_cast<float>(c[i]);
      std::memcpy(&bytes[i * elem], &v, elem);
    } else {
      std::memcpy(&bytes[i * elem], &c[i], elem);
    }
  }
  void* d = nullptr;
  hipError_t e = hipMalloc(&d, bytes.size());
  if (e == hipSuccess) e = hipMemcpy(d, bytes.data(), bytes.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (d) (void)hipFree(d);
    return tg::hip_fail(e, "pull-back coefficients");
  }
  g->pull_back_coef.emplace(key, d);
  *out = d;
  return TG_OK;
}

template <typename T>
int launch(tg_geometry* g, const tg::Block& b, const void* coef, const void* x, uint64_t nq,
           uint32_t tile, uint32_t max_iterations, double tolerance, uint32_t options, void* xi,
           void* step, hipStream_t s) {
  tg::PullBackParams<T> p{};
  p.num_cells = b.num_cells;
  p.padded_cells = (b.num_cells + tile - 1) / tile * tile;
  p.last_node = static_cast<uint32_t>(g->num_nodes - 1);
  p.nq = static_cast<uint32_t>(nq);
  uint32_t tl = 0;
  while ((1u << tl) < tile) ++tl;
  p.tile_log2 = tl;
  p.max_iterations = max_iterations;
  p.shift = (sizeof(T) == 4 && !(options & TG_OPT_NO_CELL_SHIFT)) ? 1u : 0u;
  // pull_back.rs:49-51, 642-644: 1e-13 in f64 and 1e-5 in f32 unless given.
  p.tol = static_cast<T>(tolerance >= 0.0 ? tolerance : (sizeof(T) == 8 ? 1e-13 : 1e-5));
  for (int i = 0; i < 3; ++i) p.origin[i] = static_cast<T>(i < g->gdim ? g->origin[i] : 0.0);
  // The vertex centroid rounded to f32, as pull_back.rs:100-104 returns it.
  const int td = tg::tdim(b.cell), nv = tg::num_vertices(b.cell);
  for (int d = 0; d < 3; ++d) {
    double acc = 0.0;
    if (d < td)
      for (int v = 0; v < nv; ++v) acc += tg::vertex(b.cell, v)[d];
    p.centroid[d] = static_cast<T>(static_cast<float>(d < td ? acc / nv : 0.0));
  }
  hipError_t e = tg::launch_pull_back<T>(
      b.cell, g->gdim, static_cast<int>(b.degree), static_cast<const T*>(g->nodes), b.dofmap,
      static_cast<const T*>(coef), static_cast<const T*>(x), static_cast<T*>(xi),
      static_cast<T*>(step), p, s);
  if (e == hipErrorNotSupported) return tg::fail(TG_ERR_UNSUPPORTED_DEGREE, "UnsupportedDegree");
  if (e != hipSuccess) return tg::hip_fail(e, "pull-back kernel launch");
  return TG_OK;
}

}  // namespace

extern "C" {

int tg_pull_back_block(const tg_geometry* g, int32_t block, const void* x, uint64_t num_points,
                       uint32_t tile, uint32_t max_iterations, double tolerance,
                       uint32_t options, void* xi, void* step, void* stream) {
  using tg::fail;
  if (!g) return fail(TG_ERR_INVALID_ARGUMENT, "null geometry");
  if (block < 0 || block >= static_cast<int>(g->blocks.size()))
    return fail(TG_ERR_INVALID_ARGUMENT, "block index out of range");
  if (tile != 1 && tile != 2 && tile != 4 && tile != 8 && tile != 16)
    return fail(TG_ERR_INVALID_ARGUMENT, "tile must be 1, 2, 4, 8 or 16");
  if (num_points > 0xffffffffull) return fail(TG_ERR_TOO_LARGE, "TooLarge: too many points");
  const tg::Block& b = g->blocks[block];
  if (!tg::pull_back_supported(b.cell, g->gdim, static_cast<int>(b.degree)))
    return fail(TG_ERR_UNSUPPORTED_DEGREE,
                std::string("UnsupportedDegree: no pull-back for ") + tg::cell_name(b.cell) +
                    " of degree " + std::to_string(b.degree) + " in " + std::to_string(g->gdim) +
                    "D");
  if (b.num_cells == 0 || num_points == 0) return TG_OK;  // pull_back.rs:589-591
  if (!x || !xi || !step) return fail(TG_ERR_INVALID_ARGUMENT, "null points or outputs");
  // The coefficient cache lives in the geometry; it does not change what the geometry means.
  tg_geometry* gm = const_cast<tg_geometry*>(g);
  const void* coef = nullptr;
  int rc = coefficients(gm, b, &coef);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  return g->dtype == TG_F32
             ? launch<float>(gm, b, coef, x, num_points, tile, max_iterations, tolerance, options,
                             xi, step, s)
             : launch<double>(gm, b, coef, x, num_points, tile, max_iterations, tolerance,
                              options, xi, step, s);
}

}  // extern "C"
