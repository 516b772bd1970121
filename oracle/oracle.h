/* FP64 CPU oracle of the geometry hot path. TEST INFRASTRUCTURE ONLY: tests, the smoke check and
 * the CPU-baseline leg of the benchmark call it as a checker or as the timed CPU port; the
 * library under gridcl_amd/ never links, loads or calls it.
 *
 * It restates, in plain C, the reference's f64 host reference and CPU factor path:
 *   - reference cells and the A.5 Lagrange node order
 *     (tetraloma-refcell/src/cell.rs:53-193, lagrange.rs:163-266);
 *   - Lagrange bases in closed form (independent of the library's Vandermonde construction):
 *     Silvester's product formula on equispaced simplices, products of 1D Lagrange polynomials on
 *     intervals, quadrilaterals, hexahedra (equispaced or GLL) and prisms (triangle x interval),
 *     the closed-form rational basis of the degree-1 pyramid, and a long-double Vandermonde solve
 *     in the pyramid's rational space x^a y^b z^c/(1-z)^min(a,b) for higher pyramids
 *     (polynomials/pyramid.rs:1-20);
 *   - X, J, det J, K, normals, facet normals per cell and point with the reference kernel's
 *     formulas and operation order: sums start from the node-0 product and run in ascending
 *     node order (tetraloma-testing/src/reference.rs:117-193, 258-354;
 *     tetraloma-geometry/src/layout.rs:138-369; tetraloma/tests/common/reference.rs:8-22);
 *   - the E-vector gather with index clamping (tetraloma-geometry/src/evector.rs:197-236);
 *   - the block error metric rel_err = max|got - want| / max|want| (reference.rs:776-817);
 *   - the mesh functional of the adjoint and its FP64 gradient with respect to node
 *     coordinates (no reference counterpart; pinned by finite differences and closed forms);
 *   - the pull-back of physical points by a fixed number of Newton steps from the reference
 *     cell's centroid, with the reference's step, box clamp and freezing rule
 *     (tetraloma-geometry/src/pull_back.rs:26-180), evaluated with this oracle's own bases.
 *
 * Pinning: node layouts against the reference's lagrange_points.json fixture, sub-entity tables
 * against tables.json, and the A.5 order of every cell to degree 3 (prisms and pyramids
 * included) through the reference's A.9 rows and DOLFINx P3 arrays (node_maps.rs:11-170) applied
 * to the reference's Gmsh corpus; bases by the Kronecker property, partition of unity and
 * reproduction of the element's space (for pyramids the whole rational space
 * x^a y^b z^c/(1-z)^min(a,b), a + c <= p, b + c <= p, to degree 4, with gradients); factors by
 * affine closed forms, unit-volume sums, K J = I and the reference's doc-test examples; the
 * pull-back by round trips (the physical points of known reference points come back), by
 * affine cells (exact after one step) and by the closest point of a surface for points off
 * it. */
#ifndef GRIDCL_ORACLE_H
#define GRIDCL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int or_tdim(int cell);
/* Sub-entity `index` of dimension `dim` of the reference cell (cell.rs:53-249): its cell type
 * and its vertices in the reference's order. Returns the vertex count, or -1 past the last. */
int or_sub_entity(int cell, int dim, int index, int* type, int* verts);
size_t or_num_nodes(int cell, int degree);
/* A.5 nodes [k][tdim]; returns the node count or -1. family 0 = equispaced, 1 = GLL. */
long or_lagrange_points(int cell, int degree, int family, double* out);
/* Values [q][k] and derivatives [q][k][tdim] (either may be NULL). Returns 0 or -1. */
int or_tabulate(int cell, int degree, int family, const double* pts, size_t nq, double* values,
                double* derivs);

/* Factors of one block, cell-major outputs [c][q][LEN] (any may be NULL):
 * nodes [n][gdim] absolute f64 coordinates; dofmap [c][k] (indices clamped to num_nodes - 1);
 * flags [c] or NULL; ref_normals [q][tdim] or NULL. threads <= 0: all OpenMP threads. */
int or_eval_block(int cell, int degree, int family, int gdim, const double* nodes,
                  size_t num_nodes, const uint32_t* dofmap, size_t num_cells,
                  const uint32_t* flags, const double* pts, size_t nq, const double* ref_normals,
                  double* x, double* j, double* detj, double* pinv, double* fnormals,
                  double* normals, int threads);

/* The CPU path timed as the baseline: E-vector gather (evector.rs:197-236) into [k][g][cells]
 * then X, J and det J with the direct kernel's order into QuadField tiles of `tile` cells, all
 * in f64, OpenMP over tiles. nodes are relative to `origin`. Returns 0. */
int or_cpu_factors(int tdim, int gdim, int k, const double* nodes, size_t num_nodes,
                   const double* origin, const uint32_t* dofmap, size_t num_cells,
                   const double* values, const double* derivs, size_t nq, int tile,
                   double* evector, double* x, double* j, double* detj, int threads);

/* Functional F = sum_c sum_q w_q det J + beta sum_c sum_q w_q |J|_F^2 / det(J)^(2/3) over a
 * 3D block, in f64, and its gradient dF/dX [n][3] (grad may be NULL). Returns F. */
double or_functional(int cell, int degree, int family, const double* nodes, size_t num_nodes,
                     const uint32_t* dofmap, size_t num_cells, const double* pts,
                     const double* wts, size_t nq, double beta, double* grad, int threads);

/* Pull-back (pull_back.rs:26-180): for each cell c and point q, the reference point xi of the
 * physical point x [c][q][gdim] (absolute coordinates) by max_it Newton steps from the cell's
 * vertex centroid rounded to f32 (pull_back.rs:100-104): Delta = K(xi) (x - X(xi)), K the
 * pseudo-inverse of J; s = |Delta|_inf (non-finite components count as DBL_MAX); a point not yet
 * frozen moves to xi + Delta clamped to [-0.1f, 1.1f] per component, records s in step and
 * freezes once s <= tol; frozen points keep xi and step. Outputs xi [c][q][tdim], step [c][q].
 * Returns 0, or -1 for an unsupported element. */
int or_pull_back(int cell, int degree, int family, int gdim, const double* nodes,
                 size_t num_nodes, const uint32_t* dofmap, size_t num_cells, const double* x,
                 size_t nq, int max_it, double tol, double* xi, double* step);

/* rel_err = max|got - want| / max|want| over n values (reference.rs:776-817). */
double or_rel_err(const double* got, const double* want, size_t n);

#ifdef __cplusplus
}
#endif

#endif
