// C-ABI entry point of the block measure (tetraloma-geometry/src/measure.rs:141, measure_block).
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/tetraloma_geom.h"
#include "handles.hpp"
#include "measure_internal.hpp"

namespace {

template <typename T>
int measure(const tg_geometry* g, const tg::Block& b, tg_pointset* pts, tg::DeviceTables* tabs,
            const double* weights, uint32_t options, double* value, hipStream_t s) {
  // Where the register-resident kernel computes det J, one kernel sums each cell's points
  // straight from it and adds each 64 cells' sums (the first blocked pass; no det J field); the
  // later passes follow.
  const uint64_t sums = tg::measure_scratch_elems(b.num_cells);
  if (tg::fused_cell_measures<T>(g, b, tabs, options)) {
    void* work = nullptr;
    hipError_t e = hipMallocAsync(&work, sums * sizeof(T), s);
    if (e != hipSuccess) return tg::hip_fail(e, "measure work arrays");
    int rc =
        tg::enqueue_cell_measures<T>(g, b, pts, tabs, weights, options, static_cast<T*>(work), s);
    if (rc == TG_OK) {
      e = tg::launch_measure_sums<T>((b.num_cells + 63) / 64, static_cast<T*>(work), value, s);
      if (e != hipSuccess) rc = tg::hip_fail(e, "measure kernel launch");
    }
    const hipError_t ef = hipFreeAsync(work, s);
    if (rc == TG_OK && ef != hipSuccess) rc = tg::hip_fail(ef, "measure work arrays");
    return rc;
  }
  // Elsewhere: the det J field (tiles of 16) and the sums behind it, stream-ordered.
  const uint64_t field = (b.num_cells + 15) / 16 * 16 * pts->nq;
  const uint64_t elems = field + sums;
  void* work = nullptr;
  hipError_t e = hipMallocAsync(&work, elems * sizeof(T), s);
  if (e != hipSuccess) return tg::hip_fail(e, "measure work arrays");
  T* detj = static_cast<T*>(work);
  tg_outputs out{};
  out.detj = detj;
  int rc = tg::enqueue_block<T>(g, b, pts, tabs, TG_DETJ, 16, options, &out, s);
  if (rc == TG_OK) {
    e = tg::launch_measure<T>(detj, weights, static_cast<uint32_t>(pts->nq), b.num_cells,
                              detj + field, value, s);
    if (e != hipSuccess) rc = tg::hip_fail(e, "measure kernel launch");
  }
  const hipError_t ef = hipFreeAsync(work, s);
  if (rc == TG_OK && ef != hipSuccess) rc = tg::hip_fail(ef, "measure work arrays");
  return rc;
}

}  // namespace

extern "C" {

int tg_measure_block(const tg_geometry* g, int32_t block, tg_pointset* pts, uint32_t options,
                     double* value, void* stream) {
  using tg::fail;
  if (!g || !pts || !value) return fail(TG_ERR_INVALID_ARGUMENT, "null geometry, point set or value");
  if (block < 0 || block >= static_cast<int>(g->blocks.size()))
    return fail(TG_ERR_INVALID_ARGUMENT, "block index out of range");
  const tg::Block& b = g->blocks[block];
  if (pts->cell != b.cell)
    return fail(TG_ERR_CELL_NOT_IN_RULE,
                std::string("CellNotInRule(\"") + tg::cell_name(b.cell) + "\")");
  if (pts->nq > 0xffffffffull) return fail(TG_ERR_TOO_LARGE, "TooLarge: too many points");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (b.num_cells == 0 || pts->nq == 0) {  // measure.rs:159-161: an empty sum
    hipError_t e = g->dtype == TG_F32
                       ? tg::launch_measure<float>(nullptr, nullptr, 0, 0, nullptr, value, s)
                       : tg::launch_measure<double>(nullptr, nullptr, 0, 0, nullptr, value, s);
    return e == hipSuccess ? TG_OK : tg::hip_fail(e, "measure kernel launch");
  }
  tg::DeviceTables* tabs = nullptr;
  int rc = tg::device_tables(pts, b.degree, b.family, g->dtype, &tabs);
  if (rc) return rc;
  const double* weights = nullptr;
  rc = tg::device_weights(pts, &weights);
  if (rc) return rc;
  return g->dtype == TG_F32 ? measure<float>(g, b, pts, tabs, weights, options, value, s)
                            : measure<double>(g, b, pts, tabs, weights, options, value, s);
}

}  // extern "C"
