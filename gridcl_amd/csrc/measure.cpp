// C-ABI entry point of the block measure (tetraloma-geometry/src/measure.rs:141, measure_block),
// and the work-memory layout it shares with the allocation-free variant (measure_noalloc.hip).
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/tetraloma_geom.h"
#include "handles.hpp"
#include "measure_internal.hpp"
#include "measure_sums.hpp"

namespace tg {
namespace {

// The schedule of one measure and the work memory it needs.
struct MeasurePlan {
  bool fused = false;  // det J summed straight from the register-resident kernel
  uint64_t field = 0;  // det J field elements (unfused schedule)
  uint64_t elems = 0;  // sum elements
  size_t bytes = 0;    // total work bytes
};

template <typename T>
MeasurePlan plan_measure(const tg_geometry* g, const Block& b, const tg_pointset* pts,
                         const DeviceTables* tabs, uint32_t options) {
  MeasurePlan m;
  m.fused = fused_cell_measures<T>(g, b, tabs, options);
  if (m.fused) {
    // The first-pass partials, then the passes that ping-pong between them and the space
    // behind them.
    const uint64_t n = (b.num_cells + kSumBlock - 1) / kSumBlock;
    m.elems = n + (n + kSumBlock - 1) / kSumBlock;
  } else {
    m.field = (b.num_cells + 15) / 16 * 16 * pts->nq;
    m.elems = measure_scratch_elems(b.num_cells);
  }
  m.bytes = (m.field + m.elems) * sizeof(T);
  return m;
}

// Enqueues the measure with `work` holding m.bytes bytes.
template <typename T>
int run_measure(const tg_geometry* g, const Block& b, tg_pointset* pts, DeviceTables* tabs,
                const double* weights, uint32_t options, const MeasurePlan& m, void* work,
                double* value, hipStream_t s) {
  if (m.fused) {
    // Where the register-resident kernel computes det J, one kernel sums each cell's points
    // straight from it and adds each 64 cells' sums (the first blocked pass; no det J field);
    // the later passes follow.
    T* sums = static_cast<T*>(work);
    int rc = enqueue_cell_measures<T>(g, b, pts, tabs, weights, options, sums, s);
    if (rc != TG_OK) return rc;
    const hipError_t e = launch_measure_sums<T>((b.num_cells + 63) / 64, sums, value, s);
    return e == hipSuccess ? TG_OK : hip_fail(e, "measure kernel launch");
  }
  // Elsewhere: the det J field (tiles of 16) and the sums behind it, stream-ordered.
  T* detj = static_cast<T*>(work);
  tg_outputs out{};
  out.detj = detj;
  int rc = enqueue_block<T>(g, b, pts, tabs, TG_DETJ, 16, options, &out, s);
  if (rc != TG_OK) return rc;
  const hipError_t e = launch_measure<T>(detj, weights, static_cast<uint32_t>(pts->nq),
                                         b.num_cells, detj + m.field, value, s);
  return e == hipSuccess ? TG_OK : hip_fail(e, "measure kernel launch");
}

}  // namespace

int measure_block_with(const tg_geometry* g, int32_t block, tg_pointset* pts, uint32_t options,
                       double* value, void* work, size_t work_bytes, hipStream_t s) {
  if (!g || !pts || !value) return fail(TG_ERR_INVALID_ARGUMENT, "null geometry, point set or value");
  if (block < 0 || block >= static_cast<int>(g->blocks.size()))
    return fail(TG_ERR_INVALID_ARGUMENT, "block index out of range");
  const Block& b = g->blocks[block];
  if (pts->cell != b.cell)
    return fail(TG_ERR_CELL_NOT_IN_RULE,
                std::string("CellNotInRule(\"") + cell_name(b.cell) + "\")");
  if (pts->nq > 0xffffffffull) return fail(TG_ERR_TOO_LARGE, "TooLarge: too many points");
  if (b.num_cells == 0 || pts->nq == 0) {  // measure.rs:159-161: an empty sum
    hipError_t e = g->dtype == TG_F32
                       ? launch_measure<float>(nullptr, nullptr, 0, 0, nullptr, value, s)
                       : launch_measure<double>(nullptr, nullptr, 0, 0, nullptr, value, s);
    return e == hipSuccess ? TG_OK : hip_fail(e, "measure kernel launch");
  }
  // With caller-provided work memory nothing is allocated: the tables must be resident.
  DeviceTables* tabs = nullptr;
  int rc = work ? resident_tables(pts, b.degree, b.family, g->dtype, &tabs)
                : device_tables(pts, b.degree, b.family, g->dtype, &tabs);
  if (rc) return rc;
  const double* weights = nullptr;
  rc = work ? resident_weights(pts, &weights) : device_weights(pts, &weights);
  if (rc) return rc;
  const bool f32 = g->dtype == TG_F32;
  const MeasurePlan m = f32 ? plan_measure<float>(g, b, pts, tabs, options)
                            : plan_measure<double>(g, b, pts, tabs, options);
  bool owned = false;
  if (work == nullptr) {
    const hipError_t e = work_alloc(&work, m.bytes, s);
    if (e != hipSuccess) return hip_fail(e, "measure work arrays");
    owned = true;
  } else if (m.bytes > work_bytes) {
    return fail(TG_ERR_TOO_LARGE,
                "TooLarge: the measure of this block needs " + std::to_string(m.bytes) +
                    " bytes of work memory, more than the " + std::to_string(work_bytes) +
                    " bytes available");
  }
  rc = f32 ? run_measure<float>(g, b, pts, tabs, weights, options, m, work, value, s)
           : run_measure<double>(g, b, pts, tabs, weights, options, m, work, value, s);
  if (owned) {
    const hipError_t ef = work_free(work, s);
    if (rc == TG_OK && ef != hipSuccess) rc = hip_fail(ef, "measure work arrays");
  }
  return rc;
}

}  // namespace tg

extern "C" {

int tg_measure_block(const tg_geometry* g, int32_t block, tg_pointset* pts, uint32_t options,
                     double* value, void* stream) {
  return tg::measure_block_with(g, block, pts, options, value, nullptr, 0,
                                static_cast<hipStream_t>(stream));
}

}  // extern "C"
