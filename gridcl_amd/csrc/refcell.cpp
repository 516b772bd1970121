// Host f64 element library; the header states what each part follows in the reference.
#include "refcell.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace tg {
namespace {

// ---- Reference cells (cell.rs:53-193) ---------------------------------------------------------

const double kVerts[8][8][3] = {
    /* point */ {{0, 0, 0}},
    /* interval */ {{0, 0, 0}, {1, 0, 0}},
    /* triangle */ {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}},
    /* quadrilateral */ {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}},
    /* tetrahedron */ {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
    /* hexahedron */
    {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}},
    /* prism */ {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}},
    /* pyramid */ {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}},
};
const int kNumVerts[8] = {1, 2, 3, 4, 4, 8, 6, 5};
const int kTdim[8] = {0, 1, 2, 2, 3, 3, 3, 3};

struct Entities {
  std::vector<std::vector<int>> lists[4];
  std::vector<Cell> types[4];
};

// Sub-entities numbered by sorted vertex lists in lexicographic order (cell.rs:116-193).
Entities make_entities(Cell c) {
  Entities e;
  const int nv = kNumVerts[c];
  for (int v = 0; v < nv; ++v) {
    e.lists[0].push_back({v});
    e.types[0].push_back(kPoint);
  }
  auto add = [&](int dim, std::vector<int> l, Cell t) {
    e.lists[dim].push_back(std::move(l));
    e.types[dim].push_back(t);
  };
  switch (c) {
    case kPoint:
      break;
    case kInterval:
      add(1, {0, 1}, kInterval);
      break;
    case kTriangle:
      for (auto l : {std::vector<int>{0, 1}, {0, 2}, {1, 2}}) add(1, l, kInterval);
      add(2, {0, 1, 2}, kTriangle);
      break;
    case kQuadrilateral:
      for (auto l : {std::vector<int>{0, 1}, {0, 2}, {1, 3}, {2, 3}}) add(1, l, kInterval);
      add(2, {0, 1, 2, 3}, kQuadrilateral);
      break;
    case kTetrahedron:
      for (auto l : {std::vector<int>{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}})
        add(1, l, kInterval);
      for (auto l : {std::vector<int>{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}})
        add(2, l, kTriangle);
      add(3, {0, 1, 2, 3}, kTetrahedron);
      break;
    case kHexahedron:
      for (auto l : {std::vector<int>{0, 1}, {0, 2}, {0, 4}, {1, 3}, {1, 5}, {2, 3}, {2, 6},
                     {3, 7}, {4, 5}, {4, 6}, {5, 7}, {6, 7}})
        add(1, l, kInterval);
      for (auto l : {std::vector<int>{0, 1, 2, 3}, {0, 1, 4, 5}, {0, 2, 4, 6}, {1, 3, 5, 7},
                     {2, 3, 6, 7}, {4, 5, 6, 7}})
        add(2, l, kQuadrilateral);
      add(3, {0, 1, 2, 3, 4, 5, 6, 7}, kHexahedron);
      break;
    case kPrism:
      for (auto l : {std::vector<int>{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 4}, {2, 5}, {3, 4},
                     {3, 5}, {4, 5}})
        add(1, l, kInterval);
      add(2, {0, 1, 2}, kTriangle);
      add(2, {0, 1, 3, 4}, kQuadrilateral);
      add(2, {0, 2, 3, 5}, kQuadrilateral);
      add(2, {1, 2, 4, 5}, kQuadrilateral);
      add(2, {3, 4, 5}, kTriangle);
      add(3, {0, 1, 2, 3, 4, 5}, kPrism);
      break;
    case kPyramid:
      for (auto l :
           {std::vector<int>{0, 1}, {0, 2}, {0, 4}, {1, 3}, {1, 4}, {2, 3}, {2, 4}, {3, 4}})
        add(1, l, kInterval);
      add(2, {0, 1, 2, 3}, kQuadrilateral);
      for (auto l : {std::vector<int>{0, 1, 4}, {0, 2, 4}, {1, 3, 4}, {2, 3, 4}})
        add(2, l, kTriangle);
      add(3, {0, 1, 2, 3, 4}, kPyramid);
      break;
  }
  return e;
}

const Entities& entities(Cell c) {
  static const Entities all[8] = {
      make_entities(kPoint),       make_entities(kInterval),   make_entities(kTriangle),
      make_entities(kQuadrilateral), make_entities(kTetrahedron), make_entities(kHexahedron),
      make_entities(kPrism),       make_entities(kPyramid)};
  return all[c];
}

// ---- Jacobi polynomials (jacobi.rs:18-55): value and derivative by the three-term recurrence --

void jacobi(int n, double a, double b, double x, double* p_out, double* dp_out) {
  if (n == 0) {
    *p_out = 1.0;
    *dp_out = 0.0;
    return;
  }
  const double ab = a + b;
  double p0 = 1.0, d0 = 0.0;
  double p1 = 0.5 * (a - b) + 0.5 * (ab + 2.0) * x, d1 = 0.5 * (ab + 2.0);
  for (int k = 2; k <= n; ++k) {
    const double c = 2.0 * k + ab;
    const double den = 2.0 * k * (k + ab) * (c - 2.0);
    const double a1 = (c - 1.0) * c * (c - 2.0);
    const double a0 = (c - 1.0) * (a * a - b * b);
    const double bb = 2.0 * (k + a - 1.0) * (k + b - 1.0) * c;
    const double p2 = ((a1 * x + a0) * p1 - bb * p0) / den;
    const double d2 = ((a1 * x + a0) * d1 + a1 * p1 - bb * d0) / den;
    p0 = p1;
    d0 = d1;
    p1 = p2;
    d1 = d2;
  }
  *p_out = p1;
  *dp_out = d1;
}

bool newton_done(double delta, double x) {
  return std::fabs(delta) <= 4.0 * 2.220446049250313e-16 * std::max(std::fabs(x), 1.0);
}

// Legendre polynomial L_n(2t-1) on [0,1] and its t-derivative, all orders 0..n.
void legendre01(int n, double t, double* v, double* dv) {
  const double x = 2.0 * t - 1.0;
  v[0] = 1.0;
  dv[0] = 0.0;
  if (n >= 1) {
    v[1] = x;
    dv[1] = 2.0;
  }
  for (int k = 2; k <= n; ++k) {
    v[k] = ((2.0 * k - 1.0) * x * v[k - 1] - (k - 1.0) * v[k - 2]) / k;
    dv[k] = ((2.0 * k - 1.0) * (2.0 * v[k - 1] + x * dv[k - 1]) - (k - 1.0) * dv[k - 2]) / k;
  }
}

// Scaled Jacobi polynomials S_n(x, s) = s^n P_n^(a,0)(x / s), n = 0..N, with their partial
// derivatives in x and s, by the homogenised three-term recurrence. They are polynomials in
// (x, s), so the collapsed simplex bases built from them have no singular points.
void scaled_jacobi(int N, double a, double x, double s, double* v, double* dx, double* ds) {
  v[0] = 1.0;
  dx[0] = 0.0;
  ds[0] = 0.0;
  if (N >= 1) {
    v[1] = 0.5 * (a * s + (a + 2.0) * x);
    dx[1] = 0.5 * (a + 2.0);
    ds[1] = 0.5 * a;
  }
  for (int n = 2; n <= N; ++n) {
    const double c = 2.0 * n + a;
    const double A = (c - 1.0) * c * (c - 2.0);
    const double B = (c - 1.0) * a * a;
    const double D = 2.0 * (n + a - 1.0) * (n - 1.0) * c;
    const double den = 2.0 * n * (n + a) * (c - 2.0);
    const double l = A * x + B * s;
    v[n] = (l * v[n - 1] - D * s * s * v[n - 2]) / den;
    dx[n] = (A * v[n - 1] + l * dx[n - 1] - D * s * s * dx[n - 2]) / den;
    ds[n] = (B * v[n - 1] + l * ds[n - 1] - 2.0 * D * s * v[n - 2] - D * s * s * ds[n - 2]) / den;
  }
}

// LU with partial pivoting of the n x n row-major matrix a; returns false if singular.
bool lu_factor(std::vector<double>& a, std::vector<int>& piv, size_t n) {
  piv.resize(n);
  for (size_t k = 0; k < n; ++k) {
    size_t p = k;
    double best = std::fabs(a[k * n + k]);
    for (size_t i = k + 1; i < n; ++i) {
      const double v = std::fabs(a[i * n + k]);
      if (v > best) {
        best = v;
        p = i;
      }
    }
    if (best == 0.0) return false;
    piv[k] = static_cast<int>(p);
    if (p != k)
      for (size_t j = 0; j < n; ++j) std::swap(a[k * n + j], a[p * n + j]);
    const double inv = 1.0 / a[k * n + k];
    for (size_t i = k + 1; i < n; ++i) {
      const double f = a[i * n + k] * inv;
      a[i * n + k] = f;
      if (f != 0.0)
        for (size_t j = k + 1; j < n; ++j) a[i * n + j] -= f * a[k * n + j];
    }
  }
  return true;
}

void lu_solve(const std::vector<double>& lu, const std::vector<int>& piv, size_t n, double* b) {
  for (size_t k = 0; k < n; ++k)
    if (static_cast<size_t>(piv[k]) != k) std::swap(b[k], b[piv[k]]);
  for (size_t i = 1; i < n; ++i) {
    double s = b[i];
    for (size_t j = 0; j < i; ++j) s -= lu[i * n + j] * b[j];
    b[i] = s;
  }
  for (size_t ii = n; ii-- > 0;) {
    double s = b[ii];
    for (size_t j = ii + 1; j < n; ++j) s -= lu[ii * n + j] * b[j];
    b[ii] = s / lu[ii * n + ii];
  }
}

}  // namespace

bool valid_cell(int cell) { return cell >= 0 && cell <= 7; }
int tdim(Cell c) { return kTdim[c]; }
int num_vertices(Cell c) { return kNumVerts[c]; }
const double* vertex(Cell c, int v) { return kVerts[c][v]; }
int num_sub_entities(Cell c, int dim) {
  return dim < 0 || dim > 3 ? 0 : static_cast<int>(entities(c).lists[dim].size());
}
const std::vector<int>& sub_entity(Cell c, int dim, int i) { return entities(c).lists[dim][i]; }
Cell sub_entity_type(Cell c, int dim, int i) { return entities(c).types[dim][i]; }

// A.4 node counts (lagrange.rs:59-79).
size_t num_lagrange_nodes(Cell c, unsigned deg) {
  const size_t p = deg, a = p + 1, b = p + 2;
  switch (c) {
    case kPoint: return 1;
    case kInterval: return a;
    case kTriangle: return a * b / 2;
    case kQuadrilateral: return a * a;
    case kTetrahedron: return a * b * (p + 3) / 6;
    case kHexahedron: return a * a * a;
    case kPrism: return a * a * b / 2;
    case kPyramid: return a * b * (2 * p + 3) / 6;
  }
  return 0;
}

// Gauss-Jacobi nodes: zeros of P_m^(alpha,0) by Newton with deflation from Chebyshev guesses,
// weights 1/((1-x^2) P'^2) scaled to the exact total 1/(alpha+1) (jacobi.rs:57-100).
void gauss_jacobi(int m, double alpha, std::vector<double>& pts, std::vector<double>& wts) {
  std::vector<double> x(m);
  for (int k = 0; k < m; ++k) {
    double r = -std::cos(M_PI * (2.0 * k + 1.0) / (2.0 * m));
    if (k > 0) r = 0.5 * (r + x[k - 1]);
    for (int it = 0; it < 100; ++it) {
      double p, dp;
      jacobi(m, alpha, 0.0, r, &p, &dp);
      double defl = 0.0;
      for (int j = 0; j < k; ++j) defl += 1.0 / (r - x[j]);
      const double delta = p / (dp - defl * p);
      r -= delta;
      if (newton_done(delta, r)) break;
    }
    x[k] = r;
  }
  std::vector<double> raw(m);
  double total = 0.0;
  for (int k = 0; k < m; ++k) {
    double p, dp;
    jacobi(m, alpha, 0.0, x[k], &p, &dp);
    raw[k] = 1.0 / ((1.0 - x[k] * x[k]) * dp * dp);
    total += raw[k];
  }
  const double scale = 1.0 / ((alpha + 1.0) * total);
  pts.resize(m);
  wts.resize(m);
  for (int k = 0; k < m; ++k) {
    pts[k] = 0.5 * (1.0 + x[k]);
    wts[k] = raw[k] * scale;
  }
}

// GLL points: endpoints, zeros of P_{m-2}^(1,1) for the lower half, mirrored exactly
// (jacobi.rs:102-140).
std::vector<double> gll_points(int m) {
  const int n = m - 1;
  std::vector<double> out(m, 0.0);
  out[n] = 1.0;
  for (int i = 1; i < m / 2; ++i) {
    double x = -std::cos(M_PI * i / n);
    for (int it = 0; it < 100; ++it) {
      double p, dp;
      jacobi(n - 1, 1.0, 1.0, x, &p, &dp);
      const double delta = p / dp;
      x -= delta;
      if (newton_done(delta, x)) break;
    }
    const double u = 1.0 - 0.5 * (1.0 + x);
    out[i] = 1.0 - u;
    out[n - i] = u;
  }
  if (m % 2 == 1) out[m / 2] = 0.5;
  return out;
}

// A.5 node order (lagrange.rs:163-266): vertices, edges, faces, interior; inside a sub-entity
// the nodes follow the entity's own vertex list, x = v0 + sum_k t_k (v_axis_k - v0).
bool lagrange_points(Cell c, unsigned deg, Family fam, std::vector<double>& out) {
  out.clear();
  if (deg == 0) return false;
  const int p = static_cast<int>(deg);
  std::vector<double> line;
  if (fam == kEquispaced) {
    for (int i = 1; i < p; ++i) line.push_back(static_cast<double>(i) / p);
  } else {
    if (c != kInterval && c != kQuadrilateral && c != kHexahedron) return false;
    std::vector<double> g = gll_points(p + 1);
    line.assign(g.begin() + 1, g.begin() + p);
  }
  const int td = kTdim[c];
  auto frac = [p](int i) { return static_cast<double>(i) / p; };
  for (int dim = 0; dim <= td; ++dim) {
    for (int e = 0; e < num_sub_entities(c, dim); ++e) {
      const std::vector<int>& ent = sub_entity(c, dim, e);
      const Cell et = sub_entity_type(c, dim, e);
      auto push = [&](std::initializer_list<double> t, std::initializer_list<int> axes) {
        for (int j = 0; j < td; ++j) {
          const double v0 = kVerts[c][ent[0]][j];
          double x = v0;
          auto ti = t.begin();
          for (int a : axes) {
            x += (*ti) * (kVerts[c][ent[a]][j] - v0);
            ++ti;
          }
          out.push_back(x);
        }
      };
      switch (et) {
        case kPoint: push({}, {}); break;
        case kInterval:
          for (double t : line) push({t}, {1});
          break;
        case kTriangle:
          for (int i1 = 1; i1 < p; ++i1)
            for (int i0 = 1; i0 < p - i1; ++i0) push({frac(i0), frac(i1)}, {1, 2});
          break;
        case kQuadrilateral:
          for (double t1 : line)
            for (double t0 : line) push({t0, t1}, {1, 2});
          break;
        case kTetrahedron:
          for (int i2 = 1; i2 < p; ++i2)
            for (int i1 = 1; i1 < p - i2; ++i1)
              for (int i0 = 1; i0 < p - i2 - i1; ++i0)
                push({frac(i0), frac(i1), frac(i2)}, {1, 2, 3});
          break;
        case kHexahedron:
          for (double t2 : line)
            for (double t1 : line)
              for (double t0 : line) push({t0, t1, t2}, {1, 2, 4});
          break;
        case kPrism:
          for (int i2 = 1; i2 < p; ++i2)
            for (int i1 = 1; i1 < p; ++i1)
              for (int i0 = 1; i0 < p - i1; ++i0)
                push({frac(i0), frac(i1), frac(i2)}, {1, 2, 3});
          break;
        case kPyramid:
          for (int i2 = 1; i2 < p - 1; ++i2)
            for (int i1 = 1; i1 < p - i2; ++i1)
              for (int i0 = 1; i0 < p - i2; ++i0) {
                out.push_back(frac(i0));
                out.push_back(frac(i1));
                out.push_back(frac(i2));
              }
          break;
      }
    }
  }
  return out.size() == num_lagrange_nodes(c, deg) * static_cast<size_t>(td);
}

namespace {
// Tensor product of 1D rules, first direction fastest, each point mapped by `map`
// (quadrature.rs:134-159).
template <class Map>
void tensor_rule(const std::vector<const std::vector<double>*>& p,
                 const std::vector<const std::vector<double>*>& w, Map map,
                 std::vector<double>& pts, std::vector<double>& wts) {
  const size_t nd = p.size();
  size_t total = 1;
  for (auto* r : w) total *= r->size();
  std::vector<size_t> idx(nd, 0);
  double t[3];
  for (size_t n = 0; n < total; ++n) {
    double weight = 1.0;
    for (size_t d = 0; d < nd; ++d) {
      t[d] = (*p[d])[idx[d]];
      weight *= (*w[d])[idx[d]];
    }
    map(t, pts);
    wts.push_back(weight);
    for (size_t d = 0; d < nd; ++d) {
      if (++idx[d] < w[d]->size()) break;
      idx[d] = 0;
    }
  }
}
}  // namespace

void gauss_rule(Cell c, unsigned degree, std::vector<double>& pts, std::vector<double>& wts) {
  pts.clear();
  wts.clear();
  const int m = static_cast<int>(degree / 2 + 1);
  std::vector<double> g, gw, j1, j1w, j2, j2w;
  gauss_jacobi(m, 0.0, g, gw);
  gauss_jacobi(m, 1.0, j1, j1w);
  gauss_jacobi(m, 2.0, j2, j2w);
  switch (c) {
    case kPoint: wts.push_back(1.0); break;
    case kInterval:
      tensor_rule({&g}, {&gw}, [](const double* t, std::vector<double>& o) { o.push_back(t[0]); },
                  pts, wts);
      break;
    case kQuadrilateral:
      tensor_rule({&g, &g}, {&gw, &gw},
                  [](const double* t, std::vector<double>& o) {
                    o.push_back(t[0]);
                    o.push_back(t[1]);
                  },
                  pts, wts);
      break;
    case kHexahedron:
      tensor_rule({&g, &g, &g}, {&gw, &gw, &gw},
                  [](const double* t, std::vector<double>& o) {
                    o.insert(o.end(), {t[0], t[1], t[2]});
                  },
                  pts, wts);
      break;
    case kTriangle:
      tensor_rule({&g, &j1}, {&gw, &j1w},
                  [](const double* t, std::vector<double>& o) {
                    o.push_back(t[0] * (1.0 - t[1]));
                    o.push_back(t[1]);
                  },
                  pts, wts);
      break;
    case kTetrahedron:
      tensor_rule({&g, &j1, &j2}, {&gw, &j1w, &j2w},
                  [](const double* t, std::vector<double>& o) {
                    const double s = 1.0 - t[2];
                    o.insert(o.end(), {t[0] * (1.0 - t[1]) * s, t[1] * s, t[2]});
                  },
                  pts, wts);
      break;
    case kPrism:
      tensor_rule({&g, &j1, &g}, {&gw, &j1w, &gw},
                  [](const double* t, std::vector<double>& o) {
                    o.insert(o.end(), {t[0] * (1.0 - t[1]), t[1], t[2]});
                  },
                  pts, wts);
      break;
    case kPyramid:
      tensor_rule({&g, &g, &j2}, {&gw, &gw, &j2w},
                  [](const double* t, std::vector<double>& o) {
                    const double s = 1.0 - t[2];
                    o.insert(o.end(), {t[0] * s, t[1] * s, t[2]});
                  },
                  pts, wts);
      break;
  }
}

// Keast's 4-point rule: barycentric (a,b,b,b) and permutations, a = (5+3 sqrt5)/20,
// b = (5-sqrt5)/20, weight 1/24 each. Exact to degree 2.
void tet_rule_keast4(std::vector<double>& pts, std::vector<double>& wts) {
  const double a = 0.58541019662496845446137605030969143531609;
  const double b = 0.13819660112501051517954131656343618822797;
  pts = {b, b, b, a, b, b, b, a, b, b, b, a};
  wts.assign(4, 1.0 / 24.0);
}

// Walkington's 14-point degree-5 rule: two 4-point orbits (a,a,a,1-3a) and one 6-point orbit
// (b,b,1/2-b,1/2-b) in barycentric coordinates. Constants solved to 40 digits from the
// moment equations of the S4-invariant polynomials of degree <= 5.
void tet_rule_14(std::vector<double>& pts, std::vector<double>& wts) {
  const double a1 = 0.09273525031089122640232391373703060524520;
  const double a2 = 0.31088591926330060979734573376345783299260;
  const double b = 0.04550370412564964949188052627933943906563;
  const double w1 = 0.01224884051939365825728503424772125058718;
  const double w2 = 0.01878132095300264179986427538888105563508;
  const double w3 = 0.007091003462846911073011571353376240296271;
  pts.clear();
  wts.clear();
  for (double a : {a1, a2}) {
    const double r = 1.0 - 3.0 * a;
    // Barycentric (l0, l1, l2, l3) -> Cartesian (l1, l2, l3); l0 = r first.
    const double q[4][3] = {{a, a, a}, {r, a, a}, {a, r, a}, {a, a, r}};
    for (auto& x : q) pts.insert(pts.end(), {x[0], x[1], x[2]});
    wts.insert(wts.end(), 4, a == a1 ? w1 : w2);
  }
  const double c = 0.5 - b;
  // (l0..l3) over the six placements of {b, b, c, c}.
  const double l[6][4] = {{b, b, c, c}, {b, c, b, c}, {b, c, c, b},
                          {c, b, b, c}, {c, b, c, b}, {c, c, b, b}};
  for (auto& x : l) pts.insert(pts.end(), {x[1], x[2], x[3]});
  wts.insert(wts.end(), 6, w3);
}

// Reference facet normals (facet.rs:46-92): a normal of each facet's affine hull, oriented
// away from the cell centroid, unit length.
std::vector<double> reference_facet_normals(Cell c) {
  const int td = kTdim[c];
  std::vector<double> out;
  if (td < 1) return out;
  double mid[3] = {0, 0, 0};
  for (int v = 0; v < kNumVerts[c]; ++v)
    for (int j = 0; j < td; ++j) mid[j] += kVerts[c][v][j] / kNumVerts[c];
  for (int f = 0; f < num_sub_entities(c, td - 1); ++f) {
    const std::vector<int>& fv = sub_entity(c, td - 1, f);
    const double* v0 = kVerts[c][fv[0]];
    double n[3] = {0, 0, 0};
    if (td == 1) {
      n[0] = 1.0;
    } else if (td == 2) {
      const double* v1 = kVerts[c][fv[1]];
      n[0] = v1[1] - v0[1];
      n[1] = -(v1[0] - v0[0]);
    } else {
      const double* v1 = kVerts[c][fv[1]];
      const double* v2 = kVerts[c][fv[2]];
      const double a[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
      const double b[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
      n[0] = a[1] * b[2] - a[2] * b[1];
      n[1] = a[2] * b[0] - a[0] * b[2];
      n[2] = a[0] * b[1] - a[1] * b[0];
    }
    double fc[3] = {0, 0, 0};
    for (int v : fv)
      for (int j = 0; j < td; ++j) fc[j] += kVerts[c][v][j] / fv.size();
    double outward = 0.0, len = 0.0;
    for (int j = 0; j < td; ++j) {
      outward += n[j] * (fc[j] - mid[j]);
      len += n[j] * n[j];
    }
    len = std::sqrt(len);
    const double s = outward > 0.0 ? 1.0 / len : -1.0 / len;
    for (int j = 0; j < td; ++j) out.push_back(n[j] * s);
  }
  return out;
}

// facet.rs:93-160: xi = v0 + sum_k t_k (v_k - v0) over the facet's vertex list.
std::vector<double> facet_points_to_cell(Cell c, int facet, const std::vector<double>& pts,
                                         size_t npts) {
  const int td = kTdim[c];
  const std::vector<int>& fv = sub_entity(c, td - 1, facet);
  const int fd = td - 1;
  std::vector<double> out;
  for (size_t q = 0; q < npts; ++q) {
    for (int j = 0; j < td; ++j) {
      const double v0 = kVerts[c][fv[0]][j];
      double x = v0;
      for (int k = 0; k < fd; ++k) x += pts[q * fd + k] * (kVerts[c][fv[k + 1]][j] - v0);
      out.push_back(x);
    }
  }
  return out;
}

// ---- Lagrange elements ------------------------------------------------------------------------

bool LagrangeElement::init(Cell c, unsigned degree, Family fam, std::string* err) {
  cell_ = c;
  degree_ = degree;
  if (c == kPoint) {
    if (err) *err = "the point cell has no Lagrange geometry";
    return false;
  }
  if (!lagrange_points(c, degree, fam, nodes_)) {
    if (err)
      *err = degree == 0 ? "unsupported degree" : "unsupported node family for this cell";
    return false;
  }
  const int p = static_cast<int>(degree);
  exps_.clear();
  // Spanning sets: Legendre products of the cell's polynomial space, and for the pyramid the
  // rational space x^a y^b z^c / (1-z)^min(a,b), a+c <= p, b+c <= p (polynomials/pyramid.rs).
  for (int c2 = 0; c2 <= (tdim(c) >= 3 ? p : 0); ++c2)
    for (int b = 0; b <= (tdim(c) >= 2 ? p : 0); ++b)
      for (int a = 0; a <= p; ++a) {
        bool keep = false;
        switch (c) {
          case kInterval: case kQuadrilateral: case kHexahedron: keep = true; break;
          case kTriangle: case kTetrahedron: keep = a + b + c2 <= p; break;
          case kPrism: keep = a + b <= p; break;
          case kPyramid: keep = a + c2 <= p && b + c2 <= p; break;
          default: break;
        }
        if (keep) exps_.insert(exps_.end(), {a, b, c2});
      }
  dim_ = exps_.size() / 3;
  if (dim_ != num_lagrange_nodes(c, degree)) {
    if (err) *err = "internal: spanning set size differs from the node count";
    return false;
  }
  const int td = tdim(c);
  std::vector<double> v(dim_ * dim_);
  std::vector<double> f(4 * dim_);
  for (size_t i = 0; i < dim_; ++i) {
    spanning(&nodes_[i * td], 0, f.data());
    for (size_t j = 0; j < dim_; ++j) v[i * dim_ + j] = f[j];
  }
  std::vector<int> piv;
  if (!lu_factor(v, piv, dim_)) {
    if (err) *err = "singular Vandermonde matrix";
    return false;
  }
  // coef_ = V^{-1}: column i solves V c = e_i; phi_i = sum_j coef_[j][i] f_j.
  coef_.assign(dim_ * dim_, 0.0);
  std::vector<double> col(dim_);
  for (size_t i = 0; i < dim_; ++i) {
    std::fill(col.begin(), col.end(), 0.0);
    col[i] = 1.0;
    lu_solve(v, piv, dim_, col.data());
    for (size_t j = 0; j < dim_; ++j) coef_[j * dim_ + i] = col[j];
  }
  return true;
}

void LagrangeElement::spanning(const double* x, int nderivs, double* f) const {
  const int td = tdim(cell_);
  const int p = static_cast<int>(degree_);
  double v[3][32], dv[3][32];
  if (cell_ == kPyramid) {
    // Monomials with the rational factor; the pyramid degree is at most a few.
    const double s = 1.0 - x[2];
    for (size_t j = 0; j < dim_; ++j) {
      const int a = exps_[3 * j], b = exps_[3 * j + 1], c = exps_[3 * j + 2];
      const int n = std::min(a, b);
      const double xa = std::pow(x[0], a), yb = std::pow(x[1], b), zc = std::pow(x[2], c);
      double r = 1.0, rz = 0.0;
      if (n > 0) {
        if (s == 0.0) {
          f[j] = 0.0;
          if (nderivs)
            for (int d = 0; d < 3; ++d) f[(1 + d) * dim_ + j] = 0.0;
          continue;
        }
        r = std::pow(s, -n);
        rz = n * r / s;
      }
      f[j] = xa * yb * zc * r;
      if (nderivs) {
        const double dxa = a ? a * std::pow(x[0], a - 1) : 0.0;
        const double dyb = b ? b * std::pow(x[1], b - 1) : 0.0;
        const double dzc = c ? c * std::pow(x[2], c - 1) : 0.0;
        f[dim_ + j] = dxa * yb * zc * r;
        f[2 * dim_ + j] = xa * dyb * zc * r;
        f[3 * dim_ + j] = xa * yb * (dzc * r + zc * rz);
      }
    }
    return;
  }
  if (cell_ == kTriangle || cell_ == kTetrahedron || cell_ == kPrism) {
    simplex_spanning(x, nderivs, f);
    return;
  }
  for (int d = 0; d < td; ++d) legendre01(p, x[d], v[d], dv[d]);
  for (size_t j = 0; j < dim_; ++j) {
    const int e[3] = {exps_[3 * j], exps_[3 * j + 1], exps_[3 * j + 2]};
    double val = 1.0;
    for (int d = 0; d < td; ++d) val *= v[d][e[d]];
    f[j] = val;
    if (nderivs) {
      for (int dd = 0; dd < td; ++dd) {
        double g = 1.0;
        for (int d = 0; d < td; ++d) g *= (d == dd) ? dv[d][e[d]] : v[d][e[d]];
        f[(1 + dd) * dim_ + j] = g;
      }
    }
  }
}

// Collapsed orthogonal (Dubiner) bases: on the triangle S_i(2x + y - 1, 1 - y) P_j^(2i+1,0)(2y - 1);
// on the tetrahedron S_i(2x + y + z - 1, 1 - y - z) S_j^(2i+1)(2y + z - 1, 1 - z)
// P_k^(2i+2j+2,0)(2z - 1); on the prism the triangle basis times P_c(2z - 1). Well conditioned at
// every degree, unlike products of one-dimensional polynomials restricted to a simplex.
void LagrangeElement::simplex_spanning(const double* x, int nderivs, double* f) const {
  const int p = static_cast<int>(degree_);
  double a0[32], a0x[32], a0s[32];
  double b[32], bx[32], bs[32];
  double c[32], cx[32], cs[32];
  const bool tet = cell_ == kTetrahedron;
  const double y = x[1], z = tdim(cell_) == 3 ? x[2] : 0.0;
  const double u1 = tet ? 2.0 * x[0] + y + z - 1.0 : 2.0 * x[0] + y - 1.0;
  const double s1 = tet ? 1.0 - y - z : 1.0 - y;
  scaled_jacobi(p, 0.0, u1, s1, a0, a0x, a0s);
  for (size_t jj = 0; jj < dim_; ++jj) {
    const int i = exps_[3 * jj], j = exps_[3 * jj + 1], k = exps_[3 * jj + 2];
    double val, gx, gy, gz = 0.0;
    if (cell_ == kTetrahedron) {
      scaled_jacobi(j, 2.0 * i + 1.0, 2.0 * y + z - 1.0, 1.0 - z, b, bx, bs);
      scaled_jacobi(k, 2.0 * i + 2.0 * j + 2.0, 2.0 * z - 1.0, 1.0, c, cx, cs);
      const double A = a0[i], Au = a0x[i], As = a0s[i];
      const double B = b[j], Bu = bx[j], Bs = bs[j];
      const double C = c[k], Cz = 2.0 * cx[k];
      val = A * B * C;
      gx = 2.0 * Au * B * C;
      gy = (Au - As) * B * C + A * 2.0 * Bu * C;
      gz = (Au - As) * B * C + A * (Bu - Bs) * C + A * B * Cz;
    } else {
      // Triangle factor, then the prism's interval factor in z.
      scaled_jacobi(j, 2.0 * i + 1.0, 2.0 * y - 1.0, 1.0, b, bx, bs);
      const double A = a0[i], Au = a0x[i], As = a0s[i];
      const double B = b[j], By = 2.0 * bx[j];
      val = A * B;
      gx = 2.0 * Au * B;
      gy = (Au - As) * B + A * By;
      if (cell_ == kPrism) {
        scaled_jacobi(k, 0.0, 2.0 * z - 1.0, 1.0, c, cx, cs);
        const double C = c[k], Cz = 2.0 * cx[k];
        gz = val * Cz;
        val *= C;
        gx *= C;
        gy *= C;
      }
    }
    f[jj] = val;
    if (nderivs) {
      f[dim_ + jj] = gx;
      f[2 * dim_ + jj] = gy;
      if (tdim(cell_) == 3) f[3 * dim_ + jj] = gz;
    }
  }
}

void LagrangeElement::tabulate(const double* pts, size_t nq, int nderivs,
                               std::vector<double>& out) const {
  const int td = tdim(cell_);
  const int nd = nderivs ? 1 + td : 1;
  out.assign(static_cast<size_t>(nd) * nq * dim_, 0.0);
  std::vector<double> f(4 * dim_);
  for (size_t q = 0; q < nq; ++q) {
    spanning(&pts[q * td], nderivs, f.data());
    for (int d = 0; d < nd; ++d) {
      double* o = &out[(static_cast<size_t>(d) * nq + q) * dim_];
      const double* fd = &f[d * dim_];
      for (size_t j = 0; j < dim_; ++j) {
        const double fj = fd[j];
        if (fj == 0.0) continue;
        const double* cj = &coef_[j * dim_];
        for (size_t i = 0; i < dim_; ++i) o[i] += fj * cj[i];
      }
    }
  }
}

}  // namespace tg
