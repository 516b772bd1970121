// Internal types shared by the C-ABI host code and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tg {

// Factor bits, identical to the TG_* factor enum of the public header.
enum : uint32_t {
  kFPoints = 1u << 0,
  kFJac = 1u << 1,
  kFDetJ = 1u << 2,
  kFPinv = 1u << 3,
  kFFacetNormals = 1u << 4,
  kFNormals = 1u << 5,
};

// How a register-resident factor kernel maps work to lanes and writes its outputs:
//   kMapCell        one lane per cell, every value stored from registers (any tile);
//   kMapCellStaged  one lane per cell; X, J, det J of each point staged in LDS and written as
//                   16-byte stores for the wavefront's four tiles (tiles of 16);
//   kMapTile        one wavefront per tile of 16 cells, lanes = 16 cells x 4 points; X, J,
//                   det J staged and written as the tile's consecutive rows (tiles of 16).
// The runtime-K kernel always runs kMapCell.
enum : uint32_t { kMapCell = 0, kMapCellStaged = 1, kMapTile = 2 };

// Cells one 256-lane workgroup covers under a mapping.
inline uint32_t cells_per_workgroup(uint32_t mapping) { return mapping == kMapTile ? 64u : 256u; }

// The mapping for (dtype, tdim, gdim, k, nq, tile), the call's options and the factors it
// writes (tuning.cpp).
uint32_t choose_mapping(int dtype, int tdim, int gdim, uint32_t k, uint32_t nq, uint32_t tile,
                        uint32_t options, uint32_t factors);
// Whether per-block launches remap workgroups to XCD-contiguous ranges (default 1;
// TETRALOMA_GEOM_XCD_REMAP=0 turns it off).
uint32_t choose_xcd_remap();

// Scalars of one factor launch. Device arrays are passed as separate restrict-qualified kernel
// parameters so that the table loads can stay in the scalar cache.
template <typename T>
struct FactorParams {
  uint64_t num_cells;     // cells of the block (> 0)
  uint64_t padded_cells;  // num_cells rounded up to whole tiles; lanes past num_cells copy the last
  uint32_t last_node;     // dofmap indices are clamped to it
  uint32_t nq;            // points
  uint32_t k;             // nodes per cell (used by the runtime-K kernel)
  uint32_t tile_log2;     // log2 of the QuadField tile
  uint32_t mask;          // factor bits
  uint32_t fixup;         // 1: recompute det J in FP64 where |det J| < tau |J|_F^3
  uint32_t mapping;       // kMapCell, kMapCellStaged or kMapTile (the last two: tiles of 16)
  uint32_t xcd_remap;     // 1: workgroups of one XCD take one contiguous range of the block
  double tau;
  T origin[3];            // origin narrowed to T, added back to X
};

struct FactorOutputs {
  void* x;
  void* j;
  void* detj;
  void* pinv;
  void* fnormals;
  void* normals;
};

// One block of a mixed-mesh launch: its arrays, scalars, outputs and shape index.
template <typename T>
struct HybridBlock {
  const T* nodes;
  const uint32_t* dofmap;
  const uint32_t* flags;
  const T* tab;
  const double* tab64;
  const T* ref_normals;
  FactorOutputs out;
  FactorParams<T> p;
  uint32_t shape;
  uint32_t pad_;
};

// One work item: the cells of one workgroup (padding included) of one block.
struct WorkItem {
  uint32_t block;
  uint32_t first_cell;
};

// Largest basis table (bytes) kMapTile stages in LDS; larger tables run kMapCellStaged.
constexpr size_t kMaxLdsTableBytes = 32 * 1024;

// Shape index of (tdim, gdim, k) in the persistent mixed-mesh kernel, or -1.
int hybrid_shape(int tdim, int gdim, uint32_t k);
// Elements of one wavefront's coordinate buffer in the mixed-mesh kernel: 16 cells of up to
// max_k nodes in gdim coordinates (a multiple of 4). At most kHybridMaxNodes nodes.
uint32_t hybrid_xs_elems(uint32_t max_k, int gdim);
constexpr uint32_t kHybridMaxNodes = 27;
// Wavefronts per workgroup of the mixed-mesh kernel (a multiple of 4); an item is one tile of 16
// cells per wavefront.
constexpr uint32_t kHybridWaves = 8;
// Dynamic LDS of the mixed-mesh kernel: staging and coordinate buffers plus the largest table.
size_t hybrid_lds_bytes(size_t elem, uint32_t xs_elems, size_t max_table_bytes);
// Workgroups that are resident on the device at once for the mixed-mesh kernel.
// hipErrorInvalidValue: `lds` does not fit one workgroup on this device.
template <typename T>
hipError_t hybrid_grid(bool shift, size_t lds, uint32_t* grid);
// Launches the mixed-mesh kernel: workgroup b's items are [ranges[r], ranges[r + 1]) with r the
// XCD-contiguous index of b (grid + 1 entries).
template <typename T>
hipError_t launch_hybrid(bool shift, const HybridBlock<T>* blocks, const WorkItem* items,
                         const uint32_t* ranges, uint32_t grid, size_t lds, uint32_t xs_elems,
                         hipStream_t stream);

// Launches the factor kernel for (T, tdim, gdim, k); `shift` selects the per-cell origin shift,
// `generic` forces the runtime-K kernel. Returns hipSuccess or the launch error, and
// hipErrorNotSupported if no kernel exists for the shape.
template <typename T>
hipError_t launch_factor(int tdim, int gdim, bool shift, bool generic, const T* nodes,
                         const uint32_t* dofmap, const uint32_t* flags, const T* tab,
                         const double* tab64, const T* ref_normals, const FactorParams<T>& p,
                         const FactorOutputs& out, hipStream_t stream);

// launch_factor for FP32 arithmetic on node coordinates stored in FP64 (always with the cell
// shift): each cell's offsets from its node 0 are taken in FP64 and rounded once to FP32.
hipError_t launch_factor_f64_nodes(int tdim, int gdim, bool generic, const double* nodes,
                                   const uint32_t* dofmap, const uint32_t* flags,
                                   const float* tab, const double* tab64,
                                   const float* ref_normals, const FactorParams<float>& p,
                                   const FactorOutputs& out, hipStream_t stream);

// Sum-factorised hexahedra (tiles of 16; X, J, det J only): degree n1 - 1 in each direction
// (n1 = 2..4 nodes per direction) at an m x m x m tensor point set (m = 1..4, first coordinate
// fastest). tab1d is [2][m][n1] (1D basis values, then derivatives, at the 1D points); lex[k]
// is the lexicographic position (a n1 + b) n1 + c of A.5 node k, (a, b, c) indexing its
// (xi, eta, zeta) node coordinates; tab64 is the full f64 table [q][k][4] for the FP64 fix-up.
// Returns hipErrorNotSupported for other (n1, m).
template <typename T>
hipError_t launch_sumfact(int n1, int m, bool shift, const T* nodes, const uint32_t* dofmap,
                          const T* tab1d, const uint32_t* lex, const double* tab64,
                          const FactorParams<T>& p, const FactorOutputs& out, hipStream_t stream);

// Sum-factorised FP32 hexahedra with the second and third sums on the matrix cores
// (sumfact_mfma.hip): n1 = 2..4, m = 2..4, tiles of 16; alex[l] is the A.5 index of the node at
// lexicographic position l = (a n1 + b) n1 + c (the inverse of lex); tab2d is [3][n1][64], the
// A operand of lane l for kind (phi x phi, phi' x phi, phi x phi') at b: the 2D table value
// T[(p, q) = l % 16][a = l / 16] (0 outside p, q < m, a < n1). hipErrorNotSupported for other
// (n1, m).
bool has_sumfact_mfma(int n1, int m);
hipError_t launch_sumfact_mfma(int n1, int m, bool shift, const float* nodes,
                               const uint32_t* dofmap, const float* tab1d, const uint32_t* alex,
                               const float* tab2d, const double* tab64,
                               const FactorParams<float>& p, const FactorOutputs& out,
                               hipStream_t stream);

// The matrix-core kernel on the FP64 nodes of an FP32 geometry: vertex mean, offsets and first
// sums in FP64 (tab1d the FP64 1D table), the second and third sums as products of the first
// sums' hi and lo FP32 parts. hipErrorNotSupported where launch_sumfact_mfma is.
hipError_t launch_sumfact_mfma_f64_nodes(int n1, int m, const double* nodes,
                                         const uint32_t* dofmap, const double* tab1d,
                                         const uint32_t* alex, const float* tab2d,
                                         const double* tab64, const FactorParams<float>& p,
                                         const FactorOutputs& out, hipStream_t stream);

// Sum-factorised hexahedra of an FP32 geometry on its FP64 nodes: the sums in FP64 (the
// arithmetic of the FP64 kernel) and X, J and det J rounded once to FP32 at their stores; tab1d
// is the FP64 1D table.
hipError_t launch_sumfact_f64_nodes(int n1, int m, const double* nodes, const uint32_t* dofmap,
                                    const double* tab1d, const uint32_t* lex, const double* tab64,
                                    const FactorParams<float>& p, const FactorOutputs& out,
                                    hipStream_t stream);

// The cooperative strategy (k up to 256 nodes): L lanes per cell, L the largest power of two not
// above min(32, k), leaves added by a pairwise tree. hipErrorNotSupported for other shapes.
template <typename T>
hipError_t launch_factor_coop(int tdim, int gdim, bool shift, const T* nodes,
                              const uint32_t* dofmap, const uint32_t* flags, const T* tab,
                              const double* tab64, const T* ref_normals, const FactorParams<T>& p,
                              const FactorOutputs& out, hipStream_t stream);

template <typename T>
hipError_t launch_evector(const T* nodes, const uint32_t* dofmap, uint64_t num_cells,
                          uint64_t padded, uint32_t k, int gdim, uint32_t last_node, T* out,
                          hipStream_t stream);

// Whether launch_factor has a register-resident kernel for (tdim, gdim, k, shift).
template <typename T>
bool has_direct_factor(int tdim, int gdim, uint32_t k, bool shift);

// Whether launch_measure_cells has a kernel for (tdim, gdim, k, shift).
template <typename T>
bool has_measure_cells(int tdim, int gdim, uint32_t k, bool shift);

// The first pass of the measure's 64-value sums, partials[i] = sum of the cell sums
// sum_q det J(c, q) w_q of cells 64 i .. 64 i + 63 (below p.num_cells), straight from the
// register kernel's det J (no field); hipErrorNotSupported for shapes without one.
template <typename T>
hipError_t launch_measure_cells(int tdim, int gdim, bool shift, const T* nodes,
                                const uint32_t* dofmap, const T* tab, const double* tab64,
                                const double* weights, const FactorParams<T>& p, T* partials,
                                hipStream_t stream);

}  // namespace tg
