// The opaque handles of the C ABI and the host helpers shared by its translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "refcell.hpp"

namespace tg {

struct Block {
  Cell cell;
  unsigned degree;
  Family family;
  const uint32_t* dofmap;
  uint64_t num_cells;
  const uint32_t* flags;
  uint32_t k;
  int tdim;
};

// Device copies of one element's tables at one point set: interleaved [q][k][4] (value, then
// the tdim derivatives, zero-padded) in the geometry's type, the same in f64, and the reference
// facet normals [q][tdim] in the geometry's type.
//
// Hexahedra at a tensor-product point set also get the sum-factorisation tables: the 1D basis
// values and derivatives at the 1D points [2][m][n1] in the geometry's type and the
// lexicographic position of each A.5 node [k] followed by its inverse [k] (sf_m = 0 when the
// point set is not a tensor product of one 1D set with the first coordinate fastest, or m > 4,
// or n1 > 4); FP32 tables also get the matrix-core kernel's A operands sf_mfma [3][n1][64].
struct DeviceTables {
  void* tab = nullptr;
  double* tab64 = nullptr;
  void* ref_normals = nullptr;
  void* sf_tab = nullptr;
  double* sf_tab64 = nullptr;  // the same 1D table in f64 (the FP64-node path of FP32 geometries)
  uint32_t* sf_lex = nullptr;
  void* sf_mfma = nullptr;
  int sf_n1 = 0, sf_m = 0;
  void* dense_mfma = nullptr;  // the table in matrix-core operand order (3D cells, FP32 or FP64)
  ~DeviceTables();
};

// Records the message of the last error on this thread and returns `code`.
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
// Stream-ordered work memory of the library's calls from a per-device pool that keeps freed
// blocks mapped (the measure's work arrays, the equispaced conversion's lattice field, the VTU
// writer's converted nodes): the pool holds the largest such set of blocks in use at once until
// the process ends.
hipError_t work_alloc(void** p, size_t bytes, hipStream_t s);
hipError_t work_free(void* p, hipStream_t s);
const char* cell_name(int c);

}  // namespace tg

struct tg_pointset {
  tg::Cell cell;
  int tdim;
  uint64_t nq;
  std::vector<double> points, weights, normals;
  bool has_normals = false;
  std::map<std::tuple<unsigned, int, int>, std::unique_ptr<tg::DeviceTables>> cache;
  double* weights_device = nullptr;  // uploaded on first use
  ~tg_pointset();
};

struct tg_geometry {
  int dtype;
  int gdim;
  const void* nodes;
  // FP32 geometries: the same coordinates stored in FP64 (caller-owned), which the evaluation
  // kernels gather instead of `nodes` when set (tg_geometry_set_nodes_f64).
  const double* nodes64 = nullptr;
  uint64_t num_nodes;
  double origin[3];
  std::vector<tg::Block> blocks;
  double tau = 1e-2;
  // The pull-back's coefficient matrices per (cell, degree, family), uploaded on first use.
  std::map<std::tuple<int, unsigned, int>, void*> pull_back_coef;
  // The equispaced lattices per (cell, degree) of the GLL conversion, made on first use, and
  // the lattice point of each local node (device, k entries).
  std::map<std::pair<int, unsigned>, std::unique_ptr<tg_pointset>> lattices;
  std::map<std::pair<int, unsigned>, uint32_t*> lattice_point_of;
  // The equispaced conversion's table of each node's first GLL occurrence (a function of the
  // dofmaps alone), built by the first conversion; `first_ready` is recorded on its stream.
  uint32_t* first_occurrence = nullptr;
  hipEvent_t first_ready = nullptr;
  ~tg_geometry() {
    for (auto& kv : pull_back_coef) (void)hipFree(kv.second);
    for (auto& kv : lattice_point_of) (void)hipFree(kv.second);
    if (first_ready) (void)hipEventDestroy(first_ready);
    if (first_occurrence) (void)hipFree(first_occurrence);
  }
};

namespace tg {
// The cached device tables of (degree, family, dtype), computed and uploaded on first use.
int device_tables(tg_pointset* p, unsigned degree, int family, int dtype, DeviceTables** out);
// The point weights on the device (f64), uploaded on first use.
int device_weights(tg_pointset* p, const double** out);
// The same two lookups without the upload: TG_ERR_INVALID_ARGUMENT where nothing is resident.
int resident_tables(tg_pointset* p, unsigned degree, int family, int dtype, DeviceTables** out);
int resident_weights(const tg_pointset* p, const double** out);
// Whether enqueue_block runs the block's factors on the dense matrix-core kernel.
bool takes_dense_mfma(const tg_geometry* g, const Block& b, const DeviceTables* tabs,
                      uint32_t factors, uint32_t tile, uint32_t options);
// Enqueues every factor in `factors` of block b at pts (the dispatch of tg_eval_block: the
// sum-factorised kernel where it applies, the dense matrix-core kernel for FP32 P3 tetrahedra,
// else the register-resident or runtime-K kernel). Returns a TG status.
template <typename T>
int enqueue_block(const tg_geometry* g, const Block& b, tg_pointset* pts, DeviceTables* tabs,
                  uint32_t factors, uint32_t tile, uint32_t options, const struct tg_outputs* out,
                  hipStream_t stream);
// Whether enqueue_cell_measures applies to block b: enqueue_block would compute its det J alone
// with a register-resident kernel (not the sum-factorised or runtime-K one).
template <typename T>
bool fused_cell_measures(const tg_geometry* g, const Block& b, const DeviceTables* tabs,
                         uint32_t options);
// Enqueues sums[i] = the sum over cells c = 64 i .. 64 i + 63 of block b, in order, of
// sum_q det J(c, q) w_q (w narrowed to T, each product rounded, q ascending): the cell sums and
// the first pass of 64-value sums, straight from the register-resident kernel's det J, without a
// det J field. Returns a TG status.
template <typename T>
int enqueue_cell_measures(const tg_geometry* g, const Block& b, tg_pointset* pts,
                          DeviceTables* tabs, const double* weights, uint32_t options, T* sums,
                          hipStream_t stream);
}  // namespace tg
