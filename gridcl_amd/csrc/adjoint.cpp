// C-ABI entry points of the mesh functional and its vector-Jacobian product.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "../../include/tetraloma_geom.h"
#include "adjoint_internal.hpp"
#include "handles.hpp"

struct tg_functional {
  const tg_geometry* g = nullptr;
  tg::Block block{};
  tg_pointset* pts = nullptr;
  tg::DeviceTables* tabs = nullptr;
  const double* wts = nullptr;
  double* partials = nullptr;   // one FP64 per forward workgroup
  void* contrib = nullptr;      // [k][cells][3] in the geometry's type
  uint64_t* offsets = nullptr;  // [nodes + 1]
  uint32_t* entries = nullptr;  // [cells][k]: CSR position of each (cell, local node) pair
  ~tg_functional() {
    for (void* p : {static_cast<void*>(partials), contrib,
                    static_cast<void*>(offsets), static_cast<void*>(entries)})
      if (p) (void)hipFree(p);
  }
};

namespace {

template <typename T>
int alloc(T** p, size_t n, const char* what) {
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(n, 1) * sizeof(T));
  return e == hipSuccess ? TG_OK : tg::hip_fail(e, what);
}

// Node incidence in CSR form from the block's dofmap (copied to the host once): per node the
// range of its (cell, local node) pairs, indices clamped like the kernels' gather.
int build_incidence(tg_functional* f) {
  const tg::Block& b = f->block;
  const uint64_t n_entries = b.num_cells * b.k;
  if (n_entries > 0xffffffffull) return tg::fail(TG_ERR_TOO_LARGE, "TooLarge: cells * k >= 2^32");
  std::vector<uint32_t> dofmap(n_entries);
  hipError_t e = hipMemcpy(dofmap.data(), b.dofmap, n_entries * sizeof(uint32_t),
                           hipMemcpyDeviceToHost);
  if (e != hipSuccess) return tg::hip_fail(e, "hipMemcpy(dofmap)");
  const uint64_t nn = f->g->num_nodes;
  const uint32_t last = static_cast<uint32_t>(nn - 1);
  std::vector<uint64_t> offsets(nn + 1, 0);
  for (uint32_t v : dofmap) ++offsets[(v < last ? v : last) + 1];
  for (uint64_t i = 0; i < nn; ++i) offsets[i + 1] += offsets[i];
  std::vector<uint64_t> fill(offsets.begin(), offsets.end() - 1);
  // The CSR position of each (cell, local node) pair: each node's pairs in ascending (cell,
  // local node) order. The vjp writes the pair's contribution there, so the node pass reads
  // each node's contributions as one contiguous run.
  std::vector<uint32_t> slot(n_entries);
  for (uint64_t i = 0; i < n_entries; ++i) {
    const uint32_t v = dofmap[i] < last ? dofmap[i] : last;
    slot[i] = static_cast<uint32_t>(fill[v]++);
  }
  int rc = alloc(&f->offsets, nn + 1, "hipMalloc(offsets)");
  if (!rc) rc = alloc(&f->entries, n_entries, "hipMalloc(slots)");
  if (rc) return rc;
  e = hipMemcpy(f->offsets, offsets.data(), (nn + 1) * sizeof(uint64_t), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(f->entries, slot.data(), n_entries * sizeof(uint32_t), hipMemcpyHostToDevice);
  return e == hipSuccess ? TG_OK : tg::hip_fail(e, "hipMemcpy(incidence)");
}

}  // namespace

extern "C" {

int tg_functional_create(const tg_geometry* g, int32_t block, tg_pointset* pts,
                         tg_functional** out) {
  if (!g || !pts || !out) return tg::fail(TG_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (block < 0 || block >= static_cast<int>(g->blocks.size()))
    return tg::fail(TG_ERR_INVALID_ARGUMENT, "block index out of range");
  const tg::Block& b = g->blocks[block];
  if (pts->cell != b.cell)
    return tg::fail(TG_ERR_CELL_NOT_IN_RULE,
                    std::string("CellNotInRule(\"") + tg::cell_name(b.cell) + "\")");
  if (b.tdim != 3 || g->gdim != 3)
    return tg::fail(TG_ERR_INVALID_ARGUMENT, "the functional is defined on 3D cells in 3D");
  if (!tg::functional_supported(b.k))
    return tg::fail(TG_ERR_UNSUPPORTED_DEGREE, std::string("UnsupportedDegree { cell: ") +
                                                   tg::cell_name(b.cell) + ", degree: " +
                                                   std::to_string(b.degree) + " }");
  if (g->num_nodes == 0) return tg::fail(TG_ERR_LENGTH_MISMATCH, "geometry has no nodes");
  auto f = std::make_unique<tg_functional>();
  f->g = g;
  f->block = b;
  f->pts = pts;
  int rc = tg::device_tables(pts, b.degree, b.family, g->dtype, &f->tabs);
  if (!rc) rc = tg::device_weights(pts, &f->wts);
  if (!rc) rc = alloc(&f->partials, tg::functional_partials(b.num_cells), "hipMalloc(partials)");
  if (!rc) {
    const size_t elem = g->dtype == TG_F32 ? sizeof(float) : sizeof(double);
    rc = alloc(reinterpret_cast<unsigned char**>(&f->contrib), b.num_cells * b.k * 3 * elem,
               "hipMalloc(contributions)");
  }
  if (!rc) rc = build_incidence(f.get());
  if (rc) return rc;
  *out = f.release();
  return TG_OK;
}

int tg_functional_destroy(tg_functional* f) {
  delete f;
  return TG_OK;
}

int tg_functional_forward(tg_functional* f, const void* nodes, double beta, double* value,
                          void* stream) {
  if (!f || !value) return tg::fail(TG_ERR_INVALID_ARGUMENT, "null functional or output");
  const tg::Block& b = f->block;
  const uint32_t last = static_cast<uint32_t>(f->g->num_nodes - 1);
  const void* x = nodes ? nodes : f->g->nodes;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e;
  if (b.num_cells == 0 || f->pts->nq == 0) {
    e = tg::launch_fill_zero(value, 1, s);
  } else if (f->g->dtype == TG_F32) {
    e = tg::launch_functional_forward<float>(
        static_cast<const float*>(x), b.dofmap, b.k, static_cast<const float*>(f->tabs->tab),
        f->wts, static_cast<uint32_t>(f->pts->nq), b.num_cells, last, beta, f->partials, value, s);
  } else {
    e = tg::launch_functional_forward<double>(
        static_cast<const double*>(x), b.dofmap, b.k, static_cast<const double*>(f->tabs->tab),
        f->wts, static_cast<uint32_t>(f->pts->nq), b.num_cells, last, beta, f->partials, value, s);
  }
  return e == hipSuccess ? TG_OK : tg::hip_fail(e, "functional forward");
}

int tg_functional_vjp(tg_functional* f, const void* nodes, double beta, const double* seed,
                      double* grad, void* stream) {
  if (!f || !grad) return tg::fail(TG_ERR_INVALID_ARGUMENT, "null functional or output");
  const tg::Block& b = f->block;
  const uint64_t nn = f->g->num_nodes;
  const uint32_t last = static_cast<uint32_t>(nn - 1);
  const void* x = nodes ? nodes : f->g->nodes;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e;
  if (b.num_cells == 0 || f->pts->nq == 0) {
    e = tg::launch_fill_zero(grad, nn * 3, s);
  } else if (f->g->dtype == TG_F32) {
    e = tg::launch_functional_vjp<float>(
        static_cast<const float*>(x), b.dofmap, b.k, static_cast<const float*>(f->tabs->tab),
        f->wts, static_cast<uint32_t>(f->pts->nq), b.num_cells, last, beta, seed,
        static_cast<float*>(f->contrib), f->offsets, f->entries, nn, grad, s);
  } else {
    e = tg::launch_functional_vjp<double>(
        static_cast<const double*>(x), b.dofmap, b.k, static_cast<const double*>(f->tabs->tab),
        f->wts, static_cast<uint32_t>(f->pts->nq), b.num_cells, last, beta, seed,
        static_cast<double*>(f->contrib), f->offsets, f->entries, nn, grad, s);
  }
  return e == hipSuccess ? TG_OK : tg::hip_fail(e, "functional vjp");
}

}  // extern "C"
