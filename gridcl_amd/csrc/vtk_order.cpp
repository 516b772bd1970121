// VTK's node order for its Lagrange cells (types 68-73) and the cell types written
// (tetraloma-refcell/src/node_maps/vtk.rs:1-44 describes VTK's conventions;
// tetraloma-io/src/vtk/cell_type.rs:41-60 the type table).
//
// The order is generated here entity by entity as lattice points of the internal reference cell
// (p times the A.1 coordinates; VTK's parametric cells share that geometry) and matched against
// the A.5 lattice:
// - curve: the two vertices, then the interior from x = 0 to x = 1;
// - triangle: vertices, edges 0->1, 1->2, 2->0, then the interior as a triangle of degree p - 3,
//   recursively;
// - quadrilateral and hexahedron: vertices counter-clockwise (the top square after the bottom),
//   edges running along +x, +y or +z (bottom square y = 0, x = 1, y = 1, x = 0; the top square
//   alike; then the vertical edges in corner order), hexahedron faces x = 0, x = 1, y = 0, y = 1,
//   z = 0, z = 1, and face and cell interiors lexicographic with the first free axis fastest;
// - tetrahedron: vertices, edges 0->1, 1->2, 2->0, 0->3, 1->3, 2->3, faces (0, 1, 3), (2, 3, 1),
//   (0, 3, 2), (0, 2, 1), each a recursive triangle of degree p - 3 one step inside the face with
//   its axes along face vertex 0 -> 1 and 0 -> 2, then the interior as a tetrahedron of degree
//   p - 4, recursively;
// - wedge: bottom vertices then top, edges 0->1, 1->2, 2->0 of the bottom then the top triangle,
//   the vertical edges, the triangle faces z = 0 and z = 1 (lexicographic, x fastest), the
//   quadrilateral faces y = 0 (x fastest), x + y = 1 (y increasing fastest) and x = 0 (y
//   decreasing fastest), each with z slowest, and the interior as lexicographic triangle layers;
// - pyramid: the linear pyramid only (base counter-clockwise, then the apex).
#include <cmath>
#include <map>
#include <mutex>

#include "vtk_internal.hpp"

namespace tg::vtk {

namespace {

using Lattice = std::array<int, 3>;

std::vector<Lattice> triangle(int p) {
  if (p < 0) return {};
  if (p == 0) return {{0, 0, 0}};
  std::vector<Lattice> out = {{0, 0, 0}, {p, 0, 0}, {0, p, 0}};
  for (int t = 1; t < p; ++t) out.push_back({t, 0, 0});
  for (int t = 1; t < p; ++t) out.push_back({p - t, t, 0});
  for (int t = 1; t < p; ++t) out.push_back({0, p - t, 0});
  for (const Lattice& s : triangle(p - 3)) out.push_back({s[0] + 1, s[1] + 1, 0});
  return out;
}

std::vector<Lattice> tetrahedron(int p) {
  if (p < 0) return {};
  if (p == 0) return {{0, 0, 0}};
  const Lattice v[4] = {{0, 0, 0}, {p, 0, 0}, {0, p, 0}, {0, 0, p}};
  std::vector<Lattice> out(v, v + 4);
  const int edges[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
  for (const auto& e : edges)
    for (int t = 1; t < p; ++t) {
      Lattice x;
      for (int d = 0; d < 3; ++d) x[d] = v[e[0]][d] + t * (v[e[1]][d] - v[e[0]][d]) / p;
      out.push_back(x);
    }
  const int faces[4][3] = {{0, 1, 3}, {2, 3, 1}, {0, 3, 2}, {0, 2, 1}};
  for (const auto& f : faces)
    for (const Lattice& s : triangle(p - 3)) {
      Lattice x;
      for (int d = 0; d < 3; ++d)
        x[d] = v[f[0]][d] + (s[0] + 1) * (v[f[1]][d] - v[f[0]][d]) / p +
               (s[1] + 1) * (v[f[2]][d] - v[f[0]][d]) / p;
      out.push_back(x);
    }
  for (const Lattice& s : tetrahedron(p - 4)) out.push_back({s[0] + 1, s[1] + 1, s[2] + 1});
  return out;
}

// The bottom square's edges (y = 0, x = p, y = p, x = 0) at height z, each along +x or +y.
void square_edges(int p, int z, std::vector<Lattice>& out) {
  for (int t = 1; t < p; ++t) out.push_back({t, 0, z});
  for (int t = 1; t < p; ++t) out.push_back({p, t, z});
  for (int t = 1; t < p; ++t) out.push_back({t, p, z});
  for (int t = 1; t < p; ++t) out.push_back({0, t, z});
}

std::vector<Lattice> quadrilateral(int p) {
  std::vector<Lattice> out = {{0, 0, 0}, {p, 0, 0}, {p, p, 0}, {0, p, 0}};
  square_edges(p, 0, out);
  for (int j = 1; j < p; ++j)
    for (int i = 1; i < p; ++i) out.push_back({i, j, 0});
  return out;
}

std::vector<Lattice> hexahedron(int p) {
  std::vector<Lattice> out = {{0, 0, 0}, {p, 0, 0}, {p, p, 0}, {0, p, 0},
                              {0, 0, p}, {p, 0, p}, {p, p, p}, {0, p, p}};
  square_edges(p, 0, out);
  square_edges(p, p, out);
  const int corners[4][2] = {{0, 0}, {p, 0}, {p, p}, {0, p}};
  for (const auto& c : corners)
    for (int t = 1; t < p; ++t) out.push_back({c[0], c[1], t});
  for (int x : {0, p})
    for (int k = 1; k < p; ++k)
      for (int j = 1; j < p; ++j) out.push_back({x, j, k});
  for (int y : {0, p})
    for (int k = 1; k < p; ++k)
      for (int i = 1; i < p; ++i) out.push_back({i, y, k});
  for (int z : {0, p})
    for (int j = 1; j < p; ++j)
      for (int i = 1; i < p; ++i) out.push_back({i, j, z});
  for (int k = 1; k < p; ++k)
    for (int j = 1; j < p; ++j)
      for (int i = 1; i < p; ++i) out.push_back({i, j, k});
  return out;
}

std::vector<Lattice> wedge(int p) {
  std::vector<Lattice> out = {{0, 0, 0}, {p, 0, 0}, {0, p, 0}, {0, 0, p}, {p, 0, p}, {0, p, p}};
  for (int z : {0, p}) {
    for (int t = 1; t < p; ++t) out.push_back({t, 0, z});
    for (int t = 1; t < p; ++t) out.push_back({p - t, t, z});
    for (int t = 1; t < p; ++t) out.push_back({0, p - t, z});
  }
  const int corners[3][2] = {{0, 0}, {p, 0}, {0, p}};
  for (const auto& c : corners)
    for (int t = 1; t < p; ++t) out.push_back({c[0], c[1], t});
  auto layer = [&](int z) {
    for (int j = 1; j < p; ++j)
      for (int i = 1; i + j < p; ++i) out.push_back({i, j, z});
  };
  layer(0);
  layer(p);
  for (int k = 1; k < p; ++k)
    for (int i = 1; i < p; ++i) out.push_back({i, 0, k});
  for (int k = 1; k < p; ++k)
    for (int j = 1; j < p; ++j) out.push_back({p - j, j, k});
  for (int k = 1; k < p; ++k)
    for (int t = 1; t < p; ++t) out.push_back({0, p - t, k});
  for (int k = 1; k < p; ++k) layer(k);
  return out;
}

std::vector<Lattice> generate(Cell c, int p) {
  switch (c) {
    case kPoint: return {{0, 0, 0}};
    case kInterval: {
      std::vector<Lattice> out = {{0, 0, 0}, {p, 0, 0}};
      for (int t = 1; t < p; ++t) out.push_back({t, 0, 0});
      return out;
    }
    case kTriangle: return triangle(p);
    case kQuadrilateral: return quadrilateral(p);
    case kTetrahedron: return tetrahedron(p);
    case kHexahedron: return hexahedron(p);
    case kPrism: return wedge(p);
    case kPyramid:
      if (p != 1) return {};
      return {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}};
  }
  return {};
}

std::vector<int> build_map(Cell c, unsigned p) {
  if (c == kPoint) return {0};
  const std::vector<Lattice> g = generate(c, static_cast<int>(p));
  std::vector<double> pts;
  if (g.empty() || !lagrange_points(c, p, kEquispaced, pts)) return {};
  const int td = tdim(c);
  std::map<Lattice, int> index;
  for (size_t k = 0; k < pts.size() / td; ++k) {
    Lattice x{0, 0, 0};
    for (int d = 0; d < td; ++d) x[d] = static_cast<int>(std::lround(pts[k * td + d] * p));
    index[x] = static_cast<int>(k);
  }
  if (g.size() != index.size()) return {};
  std::vector<int> map;
  for (const Lattice& x : g) {
    auto it = index.find(x);
    if (it == index.end()) return {};
    map.push_back(it->second);
  }
  return map;
}

std::mutex g_mu;

}  // namespace

const std::vector<int>& vtk_to_internal(Cell cell, unsigned p) {
  static std::map<std::pair<int, unsigned>, std::vector<int>> cache;
  std::lock_guard<std::mutex> lock(g_mu);
  auto key = std::make_pair(static_cast<int>(cell), p);
  auto it = cache.find(key);
  if (it == cache.end()) it = cache.emplace(key, build_map(cell, p)).first;
  return it->second;
}

uint8_t cell_type(Cell cell, unsigned degree, bool lagrange) {
  static const uint8_t linear[8] = {1, 3, 5, 9, 10, 12, 13, 14};
  static const uint8_t high[8] = {1, 68, 69, 70, 71, 72, 73, 14};
  return lagrange && degree >= 2 ? high[cell] : linear[cell];
}

}  // namespace tg::vtk
