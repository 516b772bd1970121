// C-ABI entry point of the GLL -> equispaced node conversion
// (tetraloma-geometry/src/equispaced.rs:158, equispaced_nodes).
#include <hip/hip_runtime.h>

#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "../../include/tetraloma_geom.h"
#include "equispaced_internal.hpp"
#include "handles.hpp"

namespace {

// The equispaced lattice of (cell, degree) as a point set, kept with the geometry so that its
// device tables outlive the launches that read them, and the lattice point of each local node.
// Tensor cells list the lattice lexicographically (first coordinate fastest), so that hexahedra
// take the sum-factorised kernel; simplices list it in A.5 order.
int lattice(tg_geometry* g, const tg::Block& b, tg_pointset** out, const uint32_t** point_of) {
  const auto key = std::make_pair(static_cast<int>(b.cell), b.degree);
  auto hit = g->lattices.find(key);
  if (hit != g->lattices.end()) {
    *out = hit->second.get();
    *point_of = g->lattice_point_of.at(key);
    return TG_OK;
  }
  auto p = std::make_unique<tg_pointset>();
  p->cell = b.cell;
  p->tdim = tg::tdim(b.cell);
  std::vector<double> a5;
  if (!tg::lagrange_points(b.cell, b.degree, tg::kEquispaced, a5))
    return tg::fail(TG_ERR_UNSUPPORTED_DEGREE, "UnsupportedDegree: no equispaced lattice");
  const int td = p->tdim;
  const size_t k = a5.size() / td;
  std::vector<uint32_t> at(k);
  const bool tensor = b.cell == tg::kInterval || b.cell == tg::kQuadrilateral ||
                      b.cell == tg::kHexahedron;
  if (tensor) {
    const unsigned n = b.degree + 1;
    p->points.assign(k * td, 0.0);
    for (size_t j = 0; j < k; ++j) {
      uint32_t lex = 0;
      for (int d = td; d-- > 0;)
        lex = lex * n + static_cast<uint32_t>(std::lround(a5[j * td + d] * b.degree));
      at[j] = lex;
      for (int d = 0; d < td; ++d) p->points[lex * td + d] = a5[j * td + d];
    }
  } else {
    p->points = a5;
    for (size_t j = 0; j < k; ++j) at[j] = static_cast<uint32_t>(j);
  }
  p->nq = k;
  p->weights.assign(p->nq, 1.0);
  uint32_t* d = nullptr;
  hipError_t e = hipMalloc(&d, k * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMemcpy(d, at.data(), k * sizeof(uint32_t), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (d) (void)hipFree(d);
    return tg::hip_fail(e, "equispaced lattice order");
  }
  *out = p.get();
  *point_of = d;
  g->lattices.emplace(key, std::move(p));
  g->lattice_point_of.emplace(key, d);
  return TG_OK;
}

struct Occurrences {
  size_t block;
  uint64_t start, len;
};

template <typename T>
int convert(tg_geometry* g, const std::vector<Occurrences>& occ, uint32_t* first, void* out,
            hipStream_t s) {
  for (const Occurrences& o : occ) {
    const tg::Block& b = g->blocks[o.block];
    tg_pointset* pts = nullptr;
    const uint32_t* point_of = nullptr;
    int rc = lattice(g, b, &pts, &point_of);
    if (rc) return rc;
    tg::DeviceTables* tabs = nullptr;
    rc = tg::device_tables(pts, b.degree, b.family, g->dtype, &tabs);
    if (rc) return rc;
    const uint64_t elems = (b.num_cells + 15) / 16 * 16 * pts->nq * g->gdim;
    void* field = nullptr;
    hipError_t e = tg::work_alloc(&field, elems * sizeof(T), s);
    if (e != hipSuccess) return tg::hip_fail(e, "equispaced conversion field");
    tg_outputs outs{};
    outs.points = field;
    // X at the lattice relative to the origin, like the node array (equispaced.rs:7-10).
    rc = tg::enqueue_block<T>(g, b, pts, tabs, TG_POINTS, 16, TG_OPT_POINTS_RELATIVE, &outs, s);
    if (rc == TG_OK) {
      e = tg::launch_select_first<T>(g->gdim, o.len, static_cast<uint32_t>(o.start), b.k,
                                     point_of, static_cast<const T*>(field), first, g->num_nodes,
                                     static_cast<T*>(out), s);
      if (e != hipSuccess) rc = tg::hip_fail(e, "equispaced conversion kernel launch");
    }
    const hipError_t ef = tg::work_free(field, s);
    if (rc) return rc;
    if (ef != hipSuccess) return tg::hip_fail(ef, "equispaced conversion field");
  }
  return TG_OK;
}

}  // namespace

extern "C" {

int tg_equispaced_nodes(const tg_geometry* g, void* out, void* stream) {
  using tg::fail;
  if (!g || !out) return fail(TG_ERR_INVALID_ARGUMENT, "null geometry or output");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t elem = g->dtype == TG_F32 ? sizeof(float) : sizeof(double);
  hipError_t e = hipMemcpyAsync(out, g->nodes, g->num_nodes * g->gdim * elem,
                                hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) return tg::hip_fail(e, "equispaced conversion copy");
  // The GLL blocks' occurrences in flat (block, cell, local node) order.
  std::vector<Occurrences> occ;
  uint64_t total = 0;
  for (size_t bi = 0; bi < g->blocks.size(); ++bi) {
    const tg::Block& b = g->blocks[bi];
    if (b.family != tg::kGll || b.num_cells == 0) continue;
    const uint64_t len = b.num_cells * b.k;
    occ.push_back({bi, total, len});
    total += len;
  }
  if (total == 0) return TG_OK;
  if (total > 0xffffffffull) return fail(TG_ERR_TOO_LARGE, "TooLarge: too many node occurrences");
  tg_geometry* gm = const_cast<tg_geometry*>(g);  // only its caches change
  if (!gm->first_occurrence) {
    // Built once: the first occurrences depend on the dofmaps only, which stay unchanged while
    // the geometry lives. Later calls, on any stream, wait for the event recorded after it.
    void* first = nullptr;
    e = hipMalloc(&first, g->num_nodes * sizeof(uint32_t));
    if (e != hipSuccess) return tg::hip_fail(e, "equispaced conversion occurrences");
    uint32_t* f = static_cast<uint32_t*>(first);
    hipEvent_t ready = nullptr;
    e = hipEventCreateWithFlags(&ready, hipEventDisableTiming);
    if (e == hipSuccess) e = tg::launch_first_init(f, g->num_nodes, s);
    for (const Occurrences& o : occ)
      if (e == hipSuccess)
        e = tg::launch_first_occurrence(g->blocks[o.block].dofmap, o.len,
                                        static_cast<uint32_t>(o.start), g->num_nodes, f, s);
    if (e == hipSuccess) e = hipEventRecord(ready, s);
    if (e != hipSuccess) {
      (void)hipStreamSynchronize(s);
      if (ready) (void)hipEventDestroy(ready);
      (void)hipFree(first);
      return tg::hip_fail(e, "equispaced conversion kernel launch");
    }
    gm->first_occurrence = f;
    gm->first_ready = ready;
  } else if (hipEventQuery(gm->first_ready) != hipSuccess) {  // still being built
    e = hipStreamWaitEvent(s, gm->first_ready, 0);
    if (e != hipSuccess) return tg::hip_fail(e, "equispaced conversion occurrences");
  }
  return g->dtype == TG_F32 ? convert<float>(gm, occ, gm->first_occurrence, out, s)
                            : convert<double>(gm, occ, gm->first_occurrence, out, s);
}

}  // extern "C"
