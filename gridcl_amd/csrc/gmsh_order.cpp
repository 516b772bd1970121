// The Gmsh element table and Gmsh's node order for Lagrange and serendipity elements.
//
// Element table: tetraloma-io/src/gmsh/element_type.rs:39-64 (the type ids of Gmsh's format,
// complete Lagrange degrees 1-10, 1-9 for hexahedra, prisms and pyramids, and the serendipity
// types 16-19).
//
// Node order (tetraloma-refcell/src/node_maps/gmsh.rs:1-35 describes Gmsh's rule): the vertices;
// the edge interiors, edge by edge from the edge's first vertex to its second; the face interiors,
// face by face, each a lower-degree Gmsh triangle (p - 3) or quadrilateral (p - 2) one lattice step
// inside the face, its u axis along face vertex 0 -> 1 and its v axis along 0 -> 2 (triangles) or
// 0 -> 3 (quadrilaterals); then the cell interior, a lower-degree Gmsh element of the same cell
// (tetrahedron p - 4, hexahedron p - 2, pyramid p - 3, triangle p - 3, quadrilateral p - 2)
// shifted one step inwards, and for the prism a Gmsh triangle of degree p - 3 (outer loop) times a
// Gmsh line of degree p - 2 (inner loop). Here the rule generates integer lattice points of the
// internal reference cell (p times the A.1 coordinates) directly, with Gmsh's vertices placed on
// the internal vertices that its reference element maps to; matching them against the A.5 lattice
// gives the permutation.
//
// Serendipity upgrade (tetraloma-refcell/src/serendipity.rs:60-281): the missing complete-P2 nodes
// are the serendipity element's interpolant at those nodes; the pyramid's only missing node, the
// base centre, uses the 8-node quadrilateral of its base.
#include <cmath>
#include <map>
#include <mutex>
#include <string>

#include "gmsh_internal.hpp"

namespace tg::gmsh {

namespace {

struct TypeRow {
  Cell cell;
  std::vector<int32_t> ids;  // degree 1 first
};

const std::vector<TypeRow>& lagrange_rows() {
  static const std::vector<TypeRow> rows = {
      {kPoint, {15}},
      {kInterval, {1, 8, 26, 27, 28, 62, 63, 64, 65, 66}},
      {kTriangle, {2, 9, 21, 23, 25, 42, 43, 44, 45, 46}},
      {kQuadrilateral, {3, 10, 36, 37, 38, 47, 48, 49, 50, 51}},
      {kTetrahedron, {4, 11, 29, 30, 31, 71, 72, 73, 74, 75}},
      {kHexahedron, {5, 12, 92, 93, 94, 95, 96, 97, 98}},
      {kPrism, {6, 13, 90, 91, 106, 107, 108, 109, 110}},
      {kPyramid, {7, 14, 118, 119, 120, 121, 122, 123, 124}},
  };
  return rows;
}

struct SerendipityRow {
  int32_t id;
  Cell cell;
  size_t num_nodes;
};

const SerendipityRow kSerendipity[] = {
    {16, kQuadrilateral, 8}, {17, kHexahedron, 20}, {18, kPrism, 15}, {19, kPyramid, 13}};

using Lattice = std::array<int, 3>;

// Gmsh's vertices on the internal lattice of degree p, in Gmsh's vertex order.
std::vector<Lattice> vertices(Cell c, int p) {
  switch (c) {
    case kPoint: return {{0, 0, 0}};
    case kInterval: return {{0, 0, 0}, {p, 0, 0}};
    case kTriangle: return {{0, 0, 0}, {p, 0, 0}, {0, p, 0}};
    case kQuadrilateral: return {{0, 0, 0}, {p, 0, 0}, {p, p, 0}, {0, p, 0}};
    case kTetrahedron: return {{0, 0, 0}, {p, 0, 0}, {0, p, 0}, {0, 0, p}};
    case kHexahedron:
      return {{0, 0, 0}, {p, 0, 0}, {p, p, 0}, {0, p, 0},
              {0, 0, p}, {p, 0, p}, {p, p, p}, {0, p, p}};
    case kPrism: return {{0, 0, 0}, {p, 0, 0}, {0, p, 0}, {0, 0, p}, {p, 0, p}, {0, p, p}};
    case kPyramid: return {{0, 0, 0}, {p, 0, 0}, {p, p, 0}, {0, p, 0}, {0, 0, p}};
  }
  return {};
}

// Gmsh's edge and face vertex lists (its MTriangle, MQuadrangle, MTetrahedron, MHexahedron,
// MPrism and MPyramid element definitions).
const std::vector<std::vector<int>>& edges(Cell c) {
  static const std::map<Cell, std::vector<std::vector<int>>> e = {
      {kInterval, {}},
      {kTriangle, {{0, 1}, {1, 2}, {2, 0}}},
      {kQuadrilateral, {{0, 1}, {1, 2}, {2, 3}, {3, 0}}},
      {kTetrahedron, {{0, 1}, {1, 2}, {2, 0}, {3, 0}, {3, 2}, {3, 1}}},
      {kHexahedron,
       {{0, 1}, {0, 3}, {0, 4}, {1, 2}, {1, 5}, {2, 3}, {2, 6}, {3, 7}, {4, 5}, {4, 7}, {5, 6},
        {6, 7}}},
      {kPrism, {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 4}, {2, 5}, {3, 4}, {3, 5}, {4, 5}}},
      {kPyramid, {{0, 1}, {0, 3}, {0, 4}, {1, 2}, {1, 4}, {2, 3}, {2, 4}, {3, 4}}},
  };
  return e.at(c);
}

const std::vector<std::vector<int>>& faces(Cell c) {
  static const std::map<Cell, std::vector<std::vector<int>>> f = {
      {kTetrahedron, {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {3, 1, 2}}},
      {kHexahedron,
       {{0, 3, 2, 1}, {0, 1, 5, 4}, {0, 4, 7, 3}, {1, 2, 6, 5}, {2, 3, 7, 6}, {4, 5, 6, 7}}},
      {kPrism, {{0, 2, 1}, {3, 4, 5}, {0, 1, 4, 3}, {0, 3, 5, 2}, {1, 2, 5, 4}}},
      {kPyramid, {{0, 1, 4}, {3, 0, 4}, {1, 2, 4}, {2, 3, 4}, {0, 3, 2, 1}}},
  };
  return f.at(c);
}

Lattice shifted(const Lattice& a, int dx, int dy, int dz) { return {a[0] + dx, a[1] + dy, a[2] + dz}; }

// The lattice points of Gmsh's degree-p element on `c` in Gmsh's order (p = 0: one point).
std::vector<Lattice> generate(Cell c, int p) {
  std::vector<Lattice> out;
  if (p == 0) return {{0, 0, 0}};
  const std::vector<Lattice> v = vertices(c, p);
  out = v;
  if (c == kPoint) return out;
  // Edges: lattice steps of (v_b - v_a) / p.
  if (c == kInterval) {
    for (int i = 1; i < p; ++i) out.push_back({i, 0, 0});
  } else {
    for (const std::vector<int>& e : edges(c))
      for (int i = 1; i < p; ++i) {
        Lattice x;
        for (int d = 0; d < 3; ++d) x[d] = v[e[0]][d] + i * (v[e[1]][d] - v[e[0]][d]) / p;
        out.push_back(x);
      }
  }
  if (c == kTriangle) {
    if (p >= 3)
      for (const Lattice& s : generate(kTriangle, p - 3)) out.push_back(shifted(s, 1, 1, 0));
    return out;
  }
  if (c == kQuadrilateral) {
    if (p >= 2)
      for (const Lattice& s : generate(kQuadrilateral, p - 2)) out.push_back(shifted(s, 1, 1, 0));
    return out;
  }
  if (c == kInterval) return out;
  // Faces.
  for (const std::vector<int>& f : faces(c)) {
    const bool tri = f.size() == 3;
    const int sub = tri ? p - 3 : p - 2;
    if (sub < 0) continue;
    const Lattice& o = v[f[0]];
    const Lattice& a = v[f[1]];
    const Lattice& b = v[f[tri ? 2 : 3]];
    for (const Lattice& s : generate(tri ? kTriangle : kQuadrilateral, sub)) {
      Lattice x;
      for (int d = 0; d < 3; ++d)
        x[d] = o[d] + (s[0] + 1) * (a[d] - o[d]) / p + (s[1] + 1) * (b[d] - o[d]) / p;
      out.push_back(x);
    }
  }
  // Interior.
  if (c == kTetrahedron && p >= 4) {
    for (const Lattice& s : generate(kTetrahedron, p - 4)) out.push_back(shifted(s, 1, 1, 1));
  } else if (c == kHexahedron && p >= 2) {
    for (const Lattice& s : generate(kHexahedron, p - 2)) out.push_back(shifted(s, 1, 1, 1));
  } else if (c == kPyramid && p >= 3) {
    for (const Lattice& s : generate(kPyramid, p - 3)) out.push_back(shifted(s, 1, 1, 1));
  } else if (c == kPrism && p >= 3) {
    const std::vector<Lattice> line = generate(kInterval, p - 2);
    for (const Lattice& t : generate(kTriangle, p - 3))
      for (const Lattice& l : line) out.push_back({t[0] + 1, t[1] + 1, l[0] + 1});
  }
  return out;
}

std::vector<int> build_map(Cell c, unsigned p) {
  const std::vector<Lattice> g = generate(c, static_cast<int>(p));
  if (c == kPoint) return {0};
  std::vector<double> pts;
  if (!lagrange_points(c, p, kEquispaced, pts)) return {};
  const int td = tdim(c);
  std::map<Lattice, int> index;
  for (size_t k = 0; k < pts.size() / td; ++k) {
    Lattice x{0, 0, 0};
    for (int d = 0; d < td; ++d) x[d] = static_cast<int>(std::lround(pts[k * td + d] * p));
    index[x] = static_cast<int>(k);
  }
  std::vector<int> map;
  if (g.size() != index.size()) return {};
  for (const Lattice& x : g) {
    auto it = index.find(x);
    if (it == index.end()) return {};
    map.push_back(it->second);
  }
  return map;
}

// Solves A x = b (n x n, row-major) by Gaussian elimination with partial pivoting.
std::vector<double> solve(std::vector<double> a, std::vector<double> b, size_t n) {
  for (size_t c = 0; c < n; ++c) {
    size_t piv = c;
    for (size_t r = c + 1; r < n; ++r)
      if (std::fabs(a[r * n + c]) > std::fabs(a[piv * n + c])) piv = r;
    if (piv != c) {
      for (size_t j = 0; j < n; ++j) std::swap(a[c * n + j], a[piv * n + j]);
      std::swap(b[c], b[piv]);
    }
    for (size_t r = c + 1; r < n; ++r) {
      const double f = a[r * n + c] / a[c * n + c];
      for (size_t j = c; j < n; ++j) a[r * n + j] -= f * a[c * n + j];
      b[r] -= f * b[c];
    }
  }
  for (size_t c = n; c-- > 0;) {
    double s = b[c];
    for (size_t j = c + 1; j < n; ++j) s -= a[c * n + j] * b[j];
    b[c] = s / a[c * n + c];
  }
  return b;
}

using Exps = std::array<int, 3>;

// The monomials of Gmsh's serendipity spaces (serendipity.rs:57-121): the quadratic
// quadrilateral's eight, the twenty of the hexahedron and the fifteen of the prism.
std::vector<Exps> serendipity_space(Cell c) {
  switch (c) {
    case kQuadrilateral:
      return {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {2, 0, 0}, {0, 2, 0}, {2, 1, 0}, {1, 2, 0}};
    case kHexahedron:
      return {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0}, {0, 2, 0}, {0, 0, 2},
              {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0},
              {0, 2, 1}, {1, 0, 2}, {0, 1, 2}, {2, 1, 1}, {1, 2, 1}, {1, 1, 2}};
    case kPrism: {
      std::vector<Exps> e;
      for (int z = 0; z <= 2; ++z)
        for (int a = 0; a <= 2; ++a)
          for (int b = 0; a + b <= 2; ++b)
            if (z < 2 || a + b <= 1) e.push_back({a, b, z});
      return e;
    }
    default: return {};
  }
}

double monomial(const Exps& e, const double* x, int td) {
  double v = 1.0;
  for (int d = 0; d < td; ++d)
    for (int k = 0; k < e[d]; ++k) v *= x[d];
  return v;
}

// Coordinates (A.1) of the degree-2 Gmsh nodes of `c`, in Gmsh order, [node][tdim].
std::vector<double> gmsh_p2_points(Cell c) {
  std::vector<double> pts;
  lagrange_points(c, 2, kEquispaced, pts);
  const int td = tdim(c);
  const std::vector<int>& map = gmsh_to_internal(c, 2);
  std::vector<double> out;
  for (int k : map)
    for (int d = 0; d < td; ++d) out.push_back(pts[k * td + d]);
  return out;
}

// The nodal basis of `space` at `nodes` ([n][td]) evaluated at x: phi solves V^T phi = m(x).
std::vector<double> nodal_at(const std::vector<Exps>& space, const std::vector<double>& nodes,
                             int td, const double* x) {
  const size_t n = space.size();
  std::vector<double> vt(n * n), m(n);
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < n; ++j) vt[j * n + i] = monomial(space[j], &nodes[i * td], td);
  for (size_t j = 0; j < n; ++j) m[j] = monomial(space[j], x, td);
  return solve(vt, m, n);
}

struct Upgrade {
  size_t nf = 0, nn = 0;
  std::vector<double> weights;
};

Upgrade build_upgrade(Cell c) {
  Upgrade u;
  for (const SerendipityRow& r : kSerendipity)
    if (r.cell == c) u.nf = r.num_nodes;
  if (u.nf == 0) return u;
  u.nn = num_lagrange_nodes(c, 2);
  const int td = tdim(c);
  const std::vector<double> pts = gmsh_p2_points(c);
  if (c != kPyramid) {
    const std::vector<Exps> space = serendipity_space(c);
    const std::vector<double> nodes(pts.begin(), pts.begin() + u.nf * td);
    for (size_t m = u.nf; m < u.nn; ++m) {
      const std::vector<double> w = nodal_at(space, nodes, td, &pts[m * td]);
      u.weights.insert(u.weights.end(), w.begin(), w.end());
    }
  } else {
    // The base (z = 0) carries the quadratic quadrilateral's eight nodes.
    std::vector<size_t> base;
    std::vector<double> nodes;
    for (size_t j = 0; j < u.nf; ++j)
      if (pts[j * 3 + 2] == 0.0) {
        base.push_back(j);
        nodes.push_back(pts[j * 3]);
        nodes.push_back(pts[j * 3 + 1]);
      }
    const std::vector<Exps> space = serendipity_space(kQuadrilateral);
    for (size_t m = u.nf; m < u.nn; ++m) {
      const std::vector<double> w = nodal_at(space, nodes, 2, &pts[m * 3]);
      std::vector<double> row(u.nf, 0.0);
      for (size_t i = 0; i < base.size(); ++i) row[base[i]] = w[i];
      u.weights.insert(u.weights.end(), row.begin(), row.end());
    }
  }
  return u;
}

std::mutex g_mu;

}  // namespace

ElementType element_type(int32_t id) {
  for (const TypeRow& r : lagrange_rows())
    for (size_t i = 0; i < r.ids.size(); ++i)
      if (r.ids[i] == id) {
        const unsigned p = static_cast<unsigned>(i + 1);
        return {id, r.cell, p, r.cell == kPoint ? 1 : num_lagrange_nodes(r.cell, p), false};
      }
  for (const SerendipityRow& r : kSerendipity)
    if (r.id == id) return {id, r.cell, 2, r.num_nodes, true};
  throw Error{16, "UnsupportedElementType: Gmsh element type " + std::to_string(id), id};
}

ElementType lagrange_element_type(Cell cell, unsigned degree) {
  for (const TypeRow& r : lagrange_rows())
    if (r.cell == cell && degree >= 1 && degree <= r.ids.size())
      return element_type(r.ids[degree - 1]);
  throw Error{4, "UnsupportedDegree: no Gmsh type for this cell of degree " + std::to_string(degree),
              degree};
}

const std::vector<int>& gmsh_to_internal(Cell cell, unsigned p) {
  static std::map<std::pair<int, unsigned>, std::vector<int>> cache;
  std::lock_guard<std::mutex> lock(g_mu);
  auto key = std::make_pair(static_cast<int>(cell), p);
  auto it = cache.find(key);
  if (it == cache.end()) it = cache.emplace(key, build_map(cell, p)).first;
  return it->second;
}

const std::vector<double>& serendipity_weights(Cell cell, size_t* num_file_nodes,
                                               size_t* num_nodes) {
  static const std::map<int, Upgrade> table = [] {
    std::map<int, Upgrade> t;
    for (const SerendipityRow& r : kSerendipity) t[r.cell] = build_upgrade(r.cell);
    return t;
  }();
  static const Upgrade none;
  auto it = table.find(static_cast<int>(cell));
  const Upgrade& u = it == table.end() ? none : it->second;
  *num_file_nodes = u.nf;
  *num_nodes = u.nn;
  return u.weights;
}

}  // namespace tg::gmsh
