// Host f64 element library: reference cells, Lagrange node layout, quadrature rules and
// tabulation of Lagrange bases. Everything here runs once per (element, point set) on the host;
// the device kernels only read the narrowed tables it produces.
//
// Follows tetraloma-refcell: cell.rs:53-249 (vertices, sub-entities), lagrange.rs:59-266 (node
// counts, A.5 node order), jacobi.rs:18-130 (Gauss-Jacobi and GLL points), quadrature.rs:77-159
// (collapsed Gauss rules), facet.rs:46-160 (reference facet normals and facet maps). The
// tabulation solves a Vandermonde system in a well-conditioned spanning set of each cell's space
// (collapsed orthogonal bases on simplices and prisms, Legendre products on tensor cells, the
// rational monomials of the pyramid); Lagrange interpolation on a node set is unique, so this
// gives the same basis as the reference's orthonormal-basis construction (tabulation.rs:78-303).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace tg {

// Cell types in the order of tetraloma_dims::CellType.
enum Cell : int {
  kPoint = 0,
  kInterval = 1,
  kTriangle = 2,
  kQuadrilateral = 3,
  kTetrahedron = 4,
  kHexahedron = 5,
  kPrism = 6,
  kPyramid = 7,
};

enum Family : int { kEquispaced = 0, kGll = 1 };

bool valid_cell(int cell);
int tdim(Cell cell);
int num_vertices(Cell cell);
// Vertex v of the reference cell, tdim coordinates.
const double* vertex(Cell cell, int v);
// Sub-entities of dimension dim: count, vertex lists and their types.
int num_sub_entities(Cell cell, int dim);
const std::vector<int>& sub_entity(Cell cell, int dim, int index);
Cell sub_entity_type(Cell cell, int dim, int index);

// Number of Lagrange nodes of degree p (A.4); 0 on overflow of a sensible range.
size_t num_lagrange_nodes(Cell cell, unsigned p);

// Lagrange nodes of degree p in A.5 order, [node][tdim]. Returns false for p == 0 or GLL on a
// cell other than interval/quadrilateral/hexahedron.
bool lagrange_points(Cell cell, unsigned p, Family family, std::vector<double>& out);

// Gauss-Jacobi rule with m points on [0,1] for the weight (1-x)^alpha; GLL points (m >= 2).
void gauss_jacobi(int m, double alpha, std::vector<double>& pts, std::vector<double>& wts);
std::vector<double> gll_points(int m);

// Collapsed Gauss rule of degree `degree` on `cell` (quadrature.rs:77): m = degree/2+1 points
// per direction, first coordinate fastest.
void gauss_rule(Cell cell, unsigned degree, std::vector<double>& pts, std::vector<double>& wts);

// Tetrahedron rules that the reference does not ship: the 4-point degree-2 rule (Keast) and the
// 14-point degree-5 rule (Walkington). Volume 1/6.
void tet_rule_keast4(std::vector<double>& pts, std::vector<double>& wts);
void tet_rule_14(std::vector<double>& pts, std::vector<double>& wts);

// Outward unit normals of the facets of `cell` in reference coordinates, [facet][tdim].
std::vector<double> reference_facet_normals(Cell cell);
// Maps facet-local points [q][tdim-1] of facet `facet` into cell coordinates [q][tdim].
std::vector<double> facet_points_to_cell(Cell cell, int facet, const std::vector<double>& pts,
                                         size_t npts);

// A Lagrange element: nodes and the inverse Vandermonde matrix of its spanning set.
class LagrangeElement {
 public:
  bool init(Cell cell, unsigned degree, Family family, std::string* err);
  size_t dim() const { return dim_; }
  Cell cell() const { return cell_; }
  const std::vector<double>& nodes() const { return nodes_; }
  // Values (nderivs = 0) or values and first derivatives (nderivs = 1) at points [q][tdim]:
  // out[((deriv)*nq + q)*dim + k], deriv 0 = value, 1+d = d/dxi_d.
  void tabulate(const double* pts, size_t nq, int nderivs, std::vector<double>& out) const;
  // The spanning set: exponents or polynomial indices (a, b, c) of function j at [3 j], and the
  // inverse Vandermonde, phi_i = sum_j coef()[j * dim + i] f_j.
  const std::vector<int>& exps() const { return exps_; }
  const std::vector<double>& coef() const { return coef_; }

 private:
  // Spanning set values and first derivatives at one point: f[(deriv)*dim + j].
  void spanning(const double* x, int nderivs, double* f) const;
  void simplex_spanning(const double* x, int nderivs, double* f) const;
  Cell cell_ = kPoint;
  unsigned degree_ = 0;
  size_t dim_ = 0;
  std::vector<double> nodes_;
  std::vector<int> exps_;     // [j][3] exponents / Legendre indices of spanning function j
  std::vector<double> coef_;  // inverse Vandermonde: phi_i = sum_j coef_[j*dim+i] f_j
};

}  // namespace tg
