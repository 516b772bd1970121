/* C ABI of the MI355X geometry library: the degree-p Lagrange geometry map of a mesh's cell
 * blocks evaluated at a point set (physical points X, Jacobians J, det J, pseudo-inverses K,
 * normals, facet normals), with the field layout, error variants and index-clamping semantics of
 * the reference's tetraloma-geometry crate.
 *
 * Conventions kept from the reference (paths relative to the reference repository):
 *   - node coordinates: [node][gdim] in the geometry's scalar type, relative to an f64 origin that
 *     the X outputs add back (tetraloma-geometry/src/geometry.rs:14-19, 139-142; layout.rs:425);
 *   - dofmaps: u32 [cell][k], node order A.5 (tetraloma-refcell/src/lagrange.rs:134-266); indices
 *     past the last node are clamped to it, not rejected (evector.rs:187-191, tests/bounds.rs);
 *   - cell flags: u32 per cell, bit 0 flips the normal of a codimension-1 cell (layout.rs:351-366);
 *   - outputs: QuadField tile layout, component k at point q of cell c at
 *       ((c / tile) * nq * LEN + q * LEN + k) * tile + c % tile,
 *     cells padded to whole tiles, padding lanes holding the last cell's values
 *     (field.rs:126-134, evector.rs:197-236). J is column-major (component d * gdim + g), K is
 *     column-major tdim x gdim (component g * tdim + t) (shape.rs; tetraloma-testing reference.rs:12-13).
 *
 * All device pointers are HIP device allocations owned by the caller unless stated otherwise.
 * Every evaluation call is asynchronous on the given stream (a hipStream_t; NULL = the null
 * stream) and synchronises nothing. Handles are thread-compatible, not thread-safe, like the
 * reference's Geometry and PointSet (tetraloma-geometry/src/testing.rs:5-6). */
#ifndef TETRALOMA_GEOM_H
#define TETRALOMA_GEOM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; one per tetraloma_device::GridError variant this path can return
 * (tetraloma-device/src/error.rs:80-150), plus argument and HIP runtime errors. */
typedef enum tg_status {
  TG_OK = 0,
  TG_ERR_LENGTH_MISMATCH = 1,      /* GridError::LengthMismatch {expected, found} */
  TG_ERR_TOO_LARGE = 2,            /* GridError::TooLarge */
  TG_ERR_ALLOCATION_TOO_LARGE = 3, /* GridError::AllocationTooLarge {bytes, max} */
  TG_ERR_UNSUPPORTED_DEGREE = 4,   /* GridError::UnsupportedDegree {cell, degree} */
  TG_ERR_UNSUPPORTED_NODE_FAMILY = 5, /* GridError::UnsupportedNodeFamily {cell, family} */
  TG_ERR_CELL_NOT_IN_RULE = 6,     /* GridError::CellNotInRule(cell) */
  TG_ERR_UNSUPPORTED_PRECISION = 7, /* GridError::UnsupportedPrecision {scalar, device} */
  TG_ERR_INVALID_ARGUMENT = 8,     /* null handle or pointer, bad tile, block index out of range */
  TG_ERR_HIP = 9,                  /* GridError::Server(..): a HIP runtime call failed */
  /* The Gmsh reader's variants (tetraloma-device/src/error.rs:129-190); tg_gmsh_last_error_detail
   * gives the number each carries. */
  TG_ERR_DIMENSION_MISMATCH = 10,  /* GridError::DimensionMismatch {expected, found} */
  TG_ERR_DUPLICATE_ORIGIN_ID = 11, /* GridError::DuplicateOriginId(id): detail = id */
  TG_ERR_DUPLICATE_CELL_BLOCK = 12, /* GridError::DuplicateCellBlock(cell) */
  TG_ERR_UNMATCHED_TAG_ELEMENT = 13, /* GridError::UnmatchedTagElement {dim, tag}: detail = tag */
  TG_ERR_IO = 14,                  /* GridError::Io(message) */
  TG_ERR_PARSE = 15,               /* GridError::Parse {format, offset, ..}: detail = offset */
  TG_ERR_UNSUPPORTED_ELEMENT_TYPE = 16, /* GridError::UnsupportedElementType: detail = type id */
  TG_ERR_INDEX_OUT_OF_RANGE = 17   /* GridError::IndexOutOfRange {index, len} (the VTU writer) */
} tg_status;

/* tetraloma_dims::CellType, in its order (tetraloma-dims/src/lib.rs). */
typedef enum tg_cell {
  TG_POINT = 0,
  TG_INTERVAL = 1,
  TG_TRIANGLE = 2,
  TG_QUADRILATERAL = 3,
  TG_TETRAHEDRON = 4,
  TG_HEXAHEDRON = 5,
  TG_PRISM = 6,
  TG_PYRAMID = 7
} tg_cell;

/* Scalar type of a geometry: its node coordinates and every field it produces. */
typedef enum tg_dtype { TG_F32 = 0, TG_F64 = 1 } tg_dtype;

/* tetraloma_refcell::NodeFamily (lagrange.rs:15-27). */
typedef enum tg_family { TG_EQUISPACED = 0, TG_GLL = 1 } tg_family;

/* Factors (layout.rs:86-101, Factor). Several may be requested in one call: they are computed in
 * one fused pass. */
enum {
  TG_POINTS = 1u << 0,        /* X, LEN gdim */
  TG_JACOBIAN = 1u << 1,      /* J, LEN tdim*gdim */
  TG_DETJ = 1u << 2,          /* det J, or sqrt(det(J^T J)) when tdim < gdim; LEN 1 */
  TG_PINV = 1u << 3,          /* K = (J^T J)^-1 J^T; LEN tdim*gdim */
  TG_FACET_NORMALS = 1u << 4, /* K^T n_ref / |K^T n_ref| (codimension 0); LEN gdim */
  TG_NORMALS = 1u << 5        /* s (J0 x J1) / |J0 x J1| (codimension 1); LEN gdim */
};

/* Evaluation options. */
enum {
  /* F32 geometries: contract node coordinates in the reference's order from node 0 instead of
   * shifting them to the cell's first node (the shift is the default). */
  TG_OPT_NO_CELL_SHIFT = 1u << 0,
  /* F32 geometries: skip the FP64 recomputation of det J at near-degenerate points. */
  TG_OPT_NO_FP64_FIXUP = 1u << 1,
  /* Force the generic runtime-K kernel (the reference's `loops` strategy) for tests. */
  TG_OPT_FORCE_GENERIC = 1u << 2,
  /* Force a lane mapping of the register-resident kernels, overriding the tuned table and the
   * TETRALOMA_GEOM_MAPPING environment variable (the reference's forced strategies,
   * tetraloma-geometry/src/strategy.rs:311-353): one lane per cell storing from registers; one
   * lane per cell with 16-byte stores staged in LDS; one wavefront per 16-cell tile streaming the
   * tile's rows. The last two need tiles of 16 (other tiles always run the first). */
  TG_OPT_MAP_CELL = 1u << 3,
  TG_OPT_MAP_CELL_STAGED = 1u << 4,
  TG_OPT_MAP_TILE = 1u << 5,
  /* Hexahedra at tensor-product point sets with tiles of 16 run the sum-factorised kernel (the
   * reference's `sumfact` strategy, tetraloma-geometry/src/sumfact/tensor.rs:337) for X, J and
   * det J where it is the faster one (FP32: degree 3, degree 2 from 3 points per direction,
   * degree 1 at 4; FP64: degree >= 2); this option keeps them on the dense kernels. */
  TG_OPT_NO_SUMFACT = 1u << 6,
  /* Plans only: one launch per block (each with the kernel tg_eval_block would run), the default;
   * it overrides TG_OPT_PLAN_PERSISTENT. tg_plan_graph_launch captures them in one graph. */
  TG_OPT_PLAN_PER_BLOCK = 1u << 7,
  /* Plans with one launch per block: capture the blocks one after another in the graph (by
   * default tg_plan_graph_launch's graph runs them as parallel branches, so they overlap;
   * tg_plan_launch always launches them in order on the caller's stream). */
  TG_OPT_PLAN_SERIAL = 1u << 8,
  /* Hexahedra at tensor-product point sets: run the sum-factorised kernel wherever it applies,
   * also for the (degree, points) pairs where the dense kernels measure faster and run by
   * default (FP32: degree 1 below 4 points per direction, degree 2 below 3; FP64: degree 1). */
  TG_OPT_SUMFACT_ALWAYS = 1u << 9,
  /* Physical points relative to the geometry's origin (the origin is not added back), like the
   * node array; the reference evaluates its equispaced conversion this way (equispaced.rs:9-10). */
  TG_OPT_POINTS_RELATIVE = 1u << 10,
  /* Plans only: run every block in the one persistent kernel over the work list, where every
   * block's element has a kernel of that form (otherwise the plan stays per block). */
  TG_OPT_PLAN_PERSISTENT = 1u << 11,
  /* Keep the node sums off the matrix cores: FP32 sum-factorised hexahedra keep their second
   * and third sums on the vector ALU (the matrix cores run them by default where they measure
   * faster: degree 2 at 3^3 points, degree 3 at 4^3 points; under TG_OPT_SUMFACT_ALWAYS
   * wherever they apply), and the 3D cells of 20 nodes and more (P3 and higher tetrahedra,
   * pyramids and prisms, Q2 and higher hexahedra off the sum-factorised point sets; FP32 and
   * FP64, at most 128 points, X, J, det J, K, facet normals) run the register-resident,
   * runtime-K or cooperative kernel instead of the dense matrix-core kernel, whose values equal
   * the register and runtime-K kernels' bit for bit. */
  TG_OPT_NO_MFMA = 1u << 12,
  /* The cooperative strategy (tetraloma-geometry/src/cooperative/factors.rs:100-183) for any
   * shape: each cell's node sums split among L lanes (L the largest power of two not above
   * min(32, k)), leaf l summing nodes l, l + L, ... from the product of node l, the leaves added
   * by a pairwise tree. It runs by default for cells of at least 56 nodes in FP32 and 84 in FP64
   * (tetrahedra from degree 5 and 6), where it measured faster than the runtime-K kernel. */
  TG_OPT_COOPERATIVE = 1u << 13
};

typedef struct tg_geometry tg_geometry;
typedef struct tg_pointset tg_pointset;

/* One cell block of a geometry (tetraloma-geometry/src/element.rs:30-70, GeometryBlock). */
typedef struct tg_block_desc {
  int32_t cell;            /* tg_cell */
  uint32_t degree;         /* Lagrange degree p >= 1 */
  int32_t family;          /* tg_family */
  const uint32_t* dofmap;  /* device, [num_cells][k], k = number of Lagrange nodes */
  uint64_t dofmap_len;     /* entries in dofmap; must be a multiple of k */
  uint32_t dofmap_stride;  /* must equal k (geometry.rs:106-150) */
  const uint32_t* flags;   /* device, [num_cells] or NULL (all zero) */
} tg_block_desc;

/* Output pointers of one evaluation; NULL for a factor not requested. Each must hold
 * tg_field_elems(num_cells, nq, LEN, tile) elements of the geometry's scalar type. */
typedef struct tg_outputs {
  void* points;
  void* jacobians;
  void* detj;
  void* pinv;
  void* facet_normals;
  void* normals;
} tg_outputs;

/* Text of the last error on this thread ("" if none). */
const char* tg_last_error_message(void);
/* Library version string. */
const char* tg_version(void);

/* Elements of a QuadField of num_cells cells, nq points, len components, in tiles of `tile`
 * cells (field.rs:234-253: cells padded to whole tiles). 0 if tile is not 1, 2, 4, 8 or 16. */
uint64_t tg_field_elems(uint64_t num_cells, uint64_t nq, uint32_t len, uint32_t tile);

/* Number of Lagrange nodes of `cell` at `degree` (tetraloma-refcell/src/lagrange.rs:59). */
uint64_t tg_num_lagrange_nodes(int32_t cell, uint32_t degree);

/* Lagrange nodes in A.5 order, [k][tdim], into `out` (host) of capacity `cap` doubles
 * (lagrange.rs:163). Writes the number of doubles to *len. */
int tg_lagrange_points(int32_t cell, uint32_t degree, int32_t family, double* out, size_t cap,
                       size_t* len);

/* Geometry over caller-owned device arrays (geometry.rs:106 Geometry::new). `nodes` is
 * [num_nodes][gdim] in `dtype`, relative to `origin` (gdim f64 values, host). The geometry keeps
 * the pointers; they must outlive it. Checks: every block's dofmap_stride equals its node count
 * and dofmap_len is a multiple of it (LengthMismatch), supported cell/degree/family. */
int tg_geometry_create(int32_t dtype, int32_t gdim, const void* nodes, uint64_t num_nodes,
                       const double* origin, const tg_block_desc* blocks, int32_t num_blocks,
                       tg_geometry** out);
int tg_geometry_destroy(tg_geometry* g);
int tg_geometry_num_blocks(const tg_geometry* g);
uint64_t tg_geometry_block_cells(const tg_geometry* g, int32_t block);
/* Threshold tau of the FP64 fix-up: points with |det J| / |J|_F^3 < tau get det J recomputed in
 * FP64 (F32 geometries). Default 1e-2. */
int tg_geometry_set_fixup_threshold(tg_geometry* g, double tau);
/* FP32 geometries: the node coordinates also stored in FP64, [num_nodes][gdim] relative to the
 * same origin (device, caller-owned, kept by pointer; NULL detaches). No counterpart in the
 * reference, whose FP32 path reads FP32 nodes (geometry.rs:106-150). With it, the evaluations of
 * tg_eval_block, the plans and tg_measure_block gather the FP64 nodes: each cell's offsets from
 * its shift point (node 0, or the vertex mean of the sum-factorised hexahedra) are taken in FP64
 * and rounded once to FP32, and the contraction stays FP32, so the FP32 rounding of node
 * coordinates far from the cell no longer enters J (it is about 1e-5 of J at h = 1/128). The
 * FP32 node array is still used by the other operations. Unsupported for FP64 geometries and
 * under TG_OPT_NO_CELL_SHIFT (the evaluations then read the FP32 nodes). */
int tg_geometry_set_nodes_f64(tg_geometry* g, const double* nodes_f64);

/* A point set of `cell` with points [nq][tdim] and weights [nq] (host f64), optionally with the
 * outward reference normal of each point's facet [nq][tdim] (a FacetPointSet,
 * facet_rule.rs:39-92). Tables are tabulated in f64 and uploaded on first use per element. */
int tg_pointset_create(int32_t cell, const double* points, const double* weights, uint64_t nq,
                       const double* facet_normals, tg_pointset** out);
/* QuadRule::gauss's point set for `cell` at `degree` (rule.rs:79; quadrature.rs:77). */
int tg_pointset_gauss(int32_t cell, uint32_t degree, tg_pointset** out);
/* FacetQuadRule::gauss's point set (facet_rule.rs:51): a Gauss rule of `degree` on every facet,
 * mapped into the cell, facet by facet, with the facets' outward reference normals. */
int tg_pointset_facet_gauss(int32_t cell, uint32_t degree, tg_pointset** out);
/* Tetrahedron rules not in the reference: rule 0 = 4-point degree 2 (Keast), 1 = 14-point
 * degree 5 (Walkington). */
int tg_pointset_tet_rule(int32_t rule, tg_pointset** out);
int tg_pointset_destroy(tg_pointset* p);
uint64_t tg_pointset_len(const tg_pointset* p);
/* Copies points ([nq][tdim]), weights ([nq]) and, for a facet point set, the reference facet
 * normals ([nq][tdim]) to host arrays (any may be NULL). Returns the number of facet-normal
 * values copied through *num_normals if that is not NULL (0 for a plain point set). */
int tg_pointset_get(const tg_pointset* p, double* points, double* weights, double* facet_normals,
                    uint64_t* num_normals);
/* Host copies of the tables of element (cell of p, degree, family) at p's points, in f64:
 * values [nq][k] and first derivatives [nq][k][tdim] (point_set.rs:312, 350). */
int tg_pointset_tabulate(const tg_pointset* p, uint32_t degree, int32_t family, double* values,
                         double* derivatives);

/* The fused factor evaluation of block `block` at the points of `pts`: every factor in
 * `factors` (TG_POINTS | TG_JACOBIAN | ...) written to the matching pointer of `out` in the
 * QuadField layout with tiles of `tile` cells. Replaces one launch_factor call per factor
 * (tetraloma-geometry/src/layout.rs:488-772) and so physical_points_block,
 * jacobians_block (points_jac.rs:42, 90), detj_block, pseudo_inverse_block (detj_pinv.rs:42,
 * 90), facet_normals_block and normals_block (facets.rs:46, 107). An empty block or point set
 * launches nothing (layout.rs:509-511). Errors: CellNotInRule if pts is of another cell type,
 * InvalidArgument for a missing output, a factor that does not apply to the block's dimensions
 * (normals need tdim = gdim - 1, facet normals tdim = gdim and a facet point set) or a bad tile. */
int tg_eval_block(const tg_geometry* g, int32_t block, tg_pointset* pts, uint32_t factors,
                  uint32_t tile, uint32_t options, const tg_outputs* out, void* stream);

/* Gathers the E-vector of a block: [k][gdim][padded cells] with padded = cells rounded up to 16,
 * padding columns copying the last cell, coordinates relative to the origin (evector.rs:197-236,
 * Geometry::evector geometry.rs:324). `out` holds k * gdim * padded scalars. */
int tg_evector_build(const tg_geometry* g, int32_t block, void* out, void* stream);
uint64_t tg_evector_elems(const tg_geometry* g, int32_t block);

/* ---- Mixed meshes: every block in one launch ---------------------------------------------------
 *
 * A plan binds, for every block b of a geometry, the point set pts[b] (of the block's cell type)
 * and the outputs outs[b] of the factors in `factors` that apply to it (normals on blocks of
 * codimension 1, facet normals on blocks of codimension 0 with a facet point set; the others on
 * every block). Replaces the reference's per-block loop over launch_factor
 * (tetraloma/src/mesh.rs:249-257 per_block; tetraloma-geometry/src/layout.rs:488). By default
 * the plan runs one launch per block with tg_eval_block's kernels: tg_plan_launch launches them
 * in order on the caller's stream; tg_plan_graph_launch's graph holds them as parallel branches
 * (captured on streams of their own, default priority, forked and joined by events;
 * TG_OPT_PLAN_SERIAL captures them in order). TG_OPT_PLAN_PERSISTENT runs one persistent kernel
 * over a work list of the items (one workgroup's cells) of all blocks instead, split on the host
 * into one contiguous range of equal estimated work per resident workgroup (the ranges of one
 * XCD adjacent), where every block's element has a kernel of that form.
 * Outputs are laid out exactly as tg_eval_block's. */
typedef struct tg_plan tg_plan;
int tg_plan_create(const tg_geometry* g, tg_pointset* const* pts, uint32_t factors,
                   uint32_t tile, uint32_t options, const tg_outputs* outs, tg_plan** out);
int tg_plan_destroy(tg_plan* plan);
/* Work items, workgroups of the persistent launch, and whether the plan is one launch. */
int tg_plan_info(const tg_plan* plan, uint32_t* num_items, uint32_t* grid, int32_t* persistent);
/* Enqueues the plan on `stream`. */
int tg_plan_launch(tg_plan* plan, void* stream);
/* Captures the plan into a hipGraph on first use (on a private stream) and replays the graph on
 * `stream`. The plan's pointers are fixed at capture: later replays write the same outputs. A
 * capture that fails releases what it made, so a later call captures afresh. */
int tg_plan_graph_launch(tg_plan* plan, void* stream);

/* ---- Measure ----------------------------------------------------------------------------------
 *
 * measure_block (tetraloma-geometry/src/measure.rs:141): the length, area or volume of block
 * `block` at `pts`, sum_c sum_q w_q det J(xi_q) (the pseudo-determinant sqrt(det(J^T J)) on a
 * curve or surface), written to *value, a device double. The arithmetic is the reference's:
 * det J as tg_eval_block computes it (tiles of 16, `options` as there), each cell's sum over its
 * points in the geometry's type (q in order from q = 0, the weights narrowed to that type;
 * measure.rs:17-41), then passes of fixed-order sums of 64 values in that type (measure.rs:44-96),
 * and the total widened to FP64. No atomics: the value is the same on every run. An empty block
 * or point set gives 0 (measure.rs:159-161). The work arrays (the det J field and the per-cell
 * sums) are allocated and freed on `stream` (hipMallocAsync / hipFreeAsync). */
int tg_measure_block(const tg_geometry* g, int32_t block, tg_pointset* pts, uint32_t options,
                     double* value, void* stream);

/* No counterpart in the reference: tg_measure_block for callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that forbid hipMalloc). Its
 * work memory is a fixed array of TG_MEASURE_SCRATCH_BYTES per device inside the library, held
 * until the process ends; nothing is allocated. It enqueues the kernels of tg_measure_block, so
 * the value is the same bit for bit. The point set's device tables and weights must be resident
 * (a tg_measure_block call with the same point set, degree, node family and precision uploads
 * them), else TG_ERR_INVALID_ARGUMENT. A block whose work memory exceeds the array is
 * TG_ERR_TOO_LARGE, before anything is enqueued: blocks measured by the register kernels need
 * ceil(cells / 64) values plus 1/64 of that, up to about 16.5 M FP32 or 8.2 M FP64 cells; the
 * others (sum-factorised, runtime-K) need the det J field, cells x points + cells values.
 * All calls on a device share its array: calls that overlap on one device (several streams or
 * host threads) race, and the caller must order them. */
#define TG_MEASURE_SCRATCH_BYTES (1u << 20)
int tg_measure_block_noalloc(const tg_geometry* g, int32_t block, tg_pointset* pts,
                             uint32_t options, double* value, void* stream);

/* ---- Pull-back --------------------------------------------------------------------------------
 *
 * pull_back_block (tetraloma-geometry/src/pull_back.rs:566): the reference points of physical
 * points. `x` is a field of block `block` (layout of tg_field_elems with LEN = gdim, tiles of
 * `tile`, num_points points per cell, absolute coordinates, the geometry's type), e.g. the output
 * of TG_POINTS; `xi` (LEN = tdim) and `step` (LEN = 1) are fields of the same tiling. Every point
 * takes `max_iterations` Newton steps (the reference's default is 16) from the reference cell's
 * vertex centroid rounded to f32: Delta = K(xi) (x - X(xi)) with K the pseudo-inverse of J
 * (least squares on curves and surfaces), xi + Delta clamped to [-0.1, 1.1] per component, and
 * |Delta|_inf (non-finite: the largest finite value) recorded in `step`; a point freezes, keeping
 * xi and step, once |Delta|_inf <= tolerance (pull_back.rs:26-180). `tolerance` < 0 takes the
 * reference's default, 1e-13 in FP64 and 1e-5 in FP32. The results are those of running every
 * step for every point (frozen points do not move; a wavefront stops once all of its points are
 * frozen), so a point's result does not depend on the others. A point with step above the
 * tolerance has not converged (its xi is an iterate). FP32 geometries work relative to each
 * cell's node 0 unless TG_OPT_NO_CELL_SHIFT. Elements: degrees 1-4 of intervals (in 1D and 2D),
 * triangles and quadrilaterals (2D and 3D), tetrahedra and prisms (3D), degrees 1-3 of
 * hexahedra and degree-1 pyramids; others are UnsupportedDegree. An empty block or point set
 * launches nothing. */
int tg_pull_back_block(const tg_geometry* g, int32_t block, const void* x, uint64_t num_points,
                       uint32_t tile, uint32_t max_iterations, double tolerance,
                       uint32_t options, void* xi, void* step, void* stream);

/* ---- GLL -> equispaced nodes ------------------------------------------------------------------
 *
 * equispaced_nodes (tetraloma-geometry/src/equispaced.rs:158): `out` ([num_nodes][gdim], the
 * geometry's type, relative to the origin like the node array) receives a copy of the nodes in
 * which every node of a GLL block holds its cell's geometry evaluated at the equispaced node of
 * the same degree (the interpolant is the same polynomial map, so the conversion is exact up to
 * rounding). A node shared by several cells takes the value of its first occurrence in (block,
 * cell, local node) order, so the result is the same on every run; nodes of equispaced blocks
 * and nodes that no GLL cell uses are copied unchanged, and dofmap entries past the node array
 * are skipped. Work arrays are allocated and freed on `stream`. The first call also builds each
 * node's first occurrence (num_nodes u32, a function of the dofmaps, which must therefore stay
 * unchanged while the geometry lives; the nodes may change) and keeps it with the geometry;
 * later calls, on any stream, reuse it. */
int tg_equispaced_nodes(const tg_geometry* g, void* out, void* stream);

/* ---- The mesh functional and its adjoint (no reference counterpart) --------------------------
 *
 * F(X) = sum_c sum_q w_q det J_cq(X) + beta sum_c sum_q w_q |J_cq(X)|_F^2 / det(J_cq(X))^(2/3)
 * over one block of 3D cells in 3D at the points and weights of a point set. The forward runs in
 * the geometry's type (F32: J of coordinates shifted to each cell's first node) and reduces in
 * FP64 in a fixed order; the vector-Jacobian product recomputes J, forms dF/dJ and each
 * (cell, local node) contribution in the geometry's type, and sums each node's contributions in
 * FP64 in CSR order (ascending cell, then local node), so both are bitwise
 * reproducible. The reduction follows Mesh::measure (tetraloma/src/mesh.rs:579;
 * tetraloma-geometry/src/measure.rs:17-102). */
typedef struct tg_functional tg_functional;

/* Binds the functional of block `block` at `pts`: builds the node incidence in CSR form (the
 * block's dofmap is copied to the host once) and allocates the work arrays (one FP64 partial
 * per 128 cells, [k][cells][3] contributions in the geometry's type). Register-resident kernels cover P1-P3
 * tetrahedra, Q1-Q2 hexahedra, P1-P2 prisms and P1 pyramids (UnsupportedDegree otherwise). */
int tg_functional_create(const tg_geometry* g, int32_t block, tg_pointset* pts,
                         tg_functional** out);
int tg_functional_destroy(tg_functional* f);
/* F into *value (a device double). `nodes` NULL: the geometry's nodes; otherwise a device array
 * of the same shape and type (the functional is then evaluated there). */
int tg_functional_forward(tg_functional* f, const void* nodes, double beta, double* value,
                          void* stream);
/* grad ([num_nodes][3] FP64, device) = seed * dF/dX, with `seed` a device double (NULL: 1). */
int tg_functional_vjp(tg_functional* f, const void* nodes, double beta, const double* seed,
                      double* grad, void* stream);

/* ---- Gmsh MSH 4.1 files (host) ----------------------------------------------------------------
 *
 * Replaces tetraloma_io::gmsh::{read_file, parse} followed by to_gmsh_mesh
 * (tetraloma-io/src/gmsh/parse.rs:571, 621; convert.rs:229): ASCII and binary files, either byte
 * order, 4- or 8-byte size_t, into the mesh the file describes. Cell blocks hold the elements of
 * the highest dimension, one block per cell type in tg_cell order, cells in file order, nodes as
 * u32 rows into the node array in A.5 order (Gmsh's node order permuted,
 * tetraloma-refcell/src/node_maps/gmsh.rs); serendipity cells (Gmsh types 16-19) are completed to
 * P2. Lower-dimensional elements become tag groups with their corner node tags. Node coordinates
 * are [node][gdim] f64 with gdim inferred from the coordinates (or the caller's, gdim > 0). A
 * cell block's `nodes` is the dofmap tg_geometry_create takes once copied to the device, with
 * family TG_EQUISPACED. Everything here runs on the host and touches no device. */
typedef struct tg_gmsh_mesh tg_gmsh_mesh;

typedef struct tg_gmsh_info {
  int32_t tdim, gdim;         /* GmshMesh::{tdim, gdim} */
  uint64_t num_nodes;         /* mesh nodes: the file's, then those the P2 completion adds */
  int32_t num_blocks, num_tag_groups, num_names;
  int32_t binary;             /* the file's encoding */
  uint64_t file_nodes;        /* nodes in $Nodes */
  int32_t num_element_blocks; /* element blocks in $Elements */
  int32_t partitioned;        /* the file has $PartitionedEntities */
  uint64_t num_partitions;
  uint64_t num_periodic;      /* links in $Periodic */
} tg_gmsh_info;

/* GmshCellBlock (convert.rs:36-52). Arrays stay owned by the mesh handle. */
typedef struct tg_gmsh_block {
  int32_t cell;              /* tg_cell */
  uint32_t degree;
  uint64_t num_cells;
  uint32_t nodes_per_cell;   /* k */
  const uint32_t* nodes;     /* [num_cells][k], A.5 order */
  const int64_t* tags;       /* element tag per cell */
  const int32_t* elementary; /* entity tag per cell */
  uint64_t num_physical;
  const int64_t* physical;   /* [num_physical][2]: (cell index, physical tag), sorted */
} tg_gmsh_block;

/* GmshTagGroup (convert.rs:82-96). */
typedef struct tg_gmsh_tag_group {
  int32_t dim;
  int32_t cell;
  uint64_t num_elements;
  uint32_t corners_per_element;
  const int64_t* corners;    /* [num_elements][vertices], node tags in internal vertex order */
  const int64_t* tags;
  const int32_t* elementary;
  uint64_t num_physical;
  const int64_t* physical;   /* [num_physical][2]: (element index, physical tag), sorted */
} tg_gmsh_tag_group;

/* One element block of $Elements as stored (model.rs, ElementBlock without its arrays). */
typedef struct tg_gmsh_element_block {
  int32_t dim, entity_tag, element_type;
  uint64_t count;
} tg_gmsh_element_block;

/* Reads the file at `path` (Io if it cannot be read) or the `len` bytes at `bytes`; gdim 0
 * infers the geometric dimension. Errors: Parse (malformed or truncated file, with the byte
 * offset), UnsupportedElementType, DuplicateOriginId, DuplicateCellBlock, DimensionMismatch,
 * UnmatchedTagElement, LengthMismatch, TooLarge, as to_gmsh_mesh (convert.rs:183-206). */
int tg_gmsh_read_file(const char* path, int32_t gdim, tg_gmsh_mesh** out);
int tg_gmsh_read_bytes(const void* bytes, uint64_t len, int32_t gdim, tg_gmsh_mesh** out);
int tg_gmsh_destroy(tg_gmsh_mesh* m);
/* The number the last reader error carries on this thread (byte offset, tag or type id). */
int64_t tg_gmsh_last_error_detail(void);
int tg_gmsh_get_info(const tg_gmsh_mesh* m, tg_gmsh_info* info);
/* [num_nodes][gdim] coordinates and the node tags (origin ids), owned by the handle. */
const double* tg_gmsh_nodes(const tg_gmsh_mesh* m);
const int64_t* tg_gmsh_node_tags(const tg_gmsh_mesh* m);
int tg_gmsh_get_block(const tg_gmsh_mesh* m, int32_t b, tg_gmsh_block* out);
int tg_gmsh_get_tag_group(const tg_gmsh_mesh* m, int32_t g, tg_gmsh_tag_group* out);
/* Physical name i ($PhysicalNames, model.rs:20-27); *name is owned by the handle. */
int tg_gmsh_get_name(const tg_gmsh_mesh* m, int32_t i, int32_t* dim, int32_t* tag,
                     const char** name);
int tg_gmsh_get_element_block(const tg_gmsh_mesh* m, int32_t i, tg_gmsh_element_block* out);
/* gmsh_to_internal (tetraloma-refcell/src/node_maps.rs:54): map[k] = the A.5 index of Gmsh's
 * k-th node of the degree-`degree` Lagrange element on `cell` (internal[map[k]] = file[k]);
 * *len = the node count (map NULL: only the count). The point cell gives {0}. */
int tg_gmsh_to_internal(int32_t cell, uint32_t degree, int32_t* map, size_t cap, size_t* len);

/* ---- VTU output (host) ------------------------------------------------------------------------
 *
 * Replaces tetraloma_io::vtk::{to_vtu, VtuFile::write_to} (tetraloma-io/src/vtk/convert.rs:116,
 * write.rs:378) and, for a device geometry on one rank, tetraloma::VtkWriter::vtu(..).write
 * (tetraloma/src/vtk.rs:736, 869): an UnstructuredGrid piece, version 2.2, little-endian,
 * UInt64 headers. Points are every node padded to three coordinates; cells of degree >= 2 are
 * VTK Lagrange cells (68-73) with their nodes in VTK's order, linear cells list their corners,
 * pyramids of degree >= 2 are written from their vertices (type 14) and counted in *pyramids. */
typedef enum tg_vtu_option {
  TG_VTU_LINEAR = 1u << 0,     /* every cell as its linear type (VtkWriter::lagrange(false)) */
  TG_VTU_BASE64 = 1u << 1,     /* base64 inside each array instead of raw appended data */
  TG_VTU_ZLIB = 1u << 2,       /* zlib-compressed data in blocks of 32 KiB */
  TG_VTU_F32_POINTS = 1u << 3  /* Float32 points (tg_vtu_write_mesh; FP32 geometries always) */
} tg_vtu_option;

typedef enum tg_vtu_type {
  TG_VTU_F32 = 0, TG_VTU_F64 = 1, TG_VTU_I32 = 2, TG_VTU_U32 = 3, TG_VTU_I64 = 4, TG_VTU_U8 = 5
} tg_vtu_type;

/* A named data array (DataArray, tetraloma-io/src/vtk/model.rs): count = entries x components. */
typedef struct tg_vtu_array {
  const char* name;
  int32_t type;         /* tg_vtu_type */
  uint32_t components;
  const void* values;   /* host for tg_vtu_write_mesh, device for tg_vtu_write_geometry */
  uint64_t count;
} tg_vtu_array;

/* One cell block (VtkCellBlock): host u32 node rows in A.5 order, len = cells x k. */
typedef struct tg_vtu_block {
  int32_t cell;
  uint32_t degree;
  const uint32_t* nodes;
  uint64_t len;
} tg_vtu_block;

/* vtk_to_internal (tetraloma-refcell/src/node_maps.rs:107): map[k] = the A.5 index of VTK's k-th
 * node (degree 1 only for the pyramid, UnsupportedDegree otherwise); map NULL: only *len. */
int tg_vtk_to_internal(int32_t cell, uint32_t degree, int32_t* map, size_t cap, size_t* len);
/* vtk_cell_type (tetraloma-io/src/vtk/cell_type.rs:41); -1 for a bad cell or degree 0. */
int32_t tg_vtk_cell_type(int32_t cell, uint32_t degree, int32_t lagrange);
/* to_vtu + write_to on host arrays: nodes [num_nodes][gdim] f64. Errors as to_vtu
 * (DimensionMismatch, LengthMismatch, UnsupportedDegree, IndexOutOfRange, TooLarge) and Io. */
int tg_vtu_write_mesh(const char* path, int32_t gdim, const double* nodes, uint64_t num_nodes,
                      const tg_vtu_block* blocks, int32_t num_blocks, uint32_t options,
                      const tg_vtu_array* point_data, int32_t num_point_data,
                      const tg_vtu_array* cell_data, int32_t num_cell_data, uint64_t* pyramids);
/* The geometry's mesh: its nodes converted to equispaced nodes on the device when a block is
 * GLL (tg_equispaced_nodes), the origin added in f64 (tetraloma/src/mesh.rs:227-241), point data
 * indexed by node and cell data by flat cell index in block order (device arrays). Synchronises
 * the stream (the file is written on return). */
int tg_vtu_write_geometry(const tg_geometry* g, const char* path, uint32_t options,
                          const tg_vtu_array* point_data, int32_t num_point_data,
                          const tg_vtu_array* cell_data, int32_t num_cell_data,
                          uint64_t* pyramids, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TETRALOMA_GEOM_H */
