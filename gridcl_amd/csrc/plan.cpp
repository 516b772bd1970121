// C-ABI entry points of the mixed-mesh evaluation plan: every block of a geometry in one
// persistent launch over a work list, optionally replayed from a captured hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/tetraloma_geom.h"
#include "geom_internal.hpp"
#include "handles.hpp"

namespace {

struct PlanBlock {
  int index;
  tg_pointset* pts;
  tg::DeviceTables* tabs;
  uint32_t mask;
  tg_outputs out;
};

}  // namespace

struct tg_plan {
  const tg_geometry* g = nullptr;
  uint32_t tile = 16, options = 0;
  bool persistent = false;
  bool shift = false;
  uint32_t grid = 0;
  uint32_t num_items = 0;
  size_t lds = 0;  // dynamic LDS of the persistent kernel
  uint32_t xs_elems = 0;  // its per-wavefront coordinate buffer
  std::vector<PlanBlock> blocks;
  void* d_blocks = nullptr;
  tg::WorkItem* d_items = nullptr;
  uint32_t* d_ranges = nullptr;  // [grid + 1]: the item range of each workgroup
  hipStream_t capture = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // Per-block plans captured in a graph: blocks 1.. run on aux[i - 1] (default priority), forked
  // from and joined to the capturing stream by events, so that the blocks' kernels overlap.
  std::vector<hipStream_t> aux;
  std::vector<hipEvent_t> events;  // [0] fork, [i] block i done
  ~tg_plan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (capture) (void)hipStreamDestroy(capture);
    for (hipStream_t a : aux) (void)hipStreamDestroy(a);
    for (hipEvent_t ev : events) (void)hipEventDestroy(ev);
    for (void* p : {d_blocks, static_cast<void*>(d_items), static_cast<void*>(d_ranges)})
      if (p) (void)hipFree(p);
  }
};

namespace {

uint32_t log2u(uint32_t v) {
  uint32_t r = 0;
  while ((1u << r) < v) ++r;
  return r;
}

template <typename T>
tg::FactorParams<T> params(const tg_geometry* g, const tg::Block& b, const tg_pointset* pts,
                           uint32_t mask, uint32_t tile, uint32_t options) {
  tg::FactorParams<T> p{};
  p.num_cells = b.num_cells;
  p.padded_cells = (b.num_cells + tile - 1) / tile * tile;
  p.last_node = static_cast<uint32_t>(g->num_nodes - 1);
  p.nq = static_cast<uint32_t>(pts->nq);
  p.k = b.k;
  p.tile_log2 = log2u(tile);
  p.mask = mask;
  const bool f32 = sizeof(T) == 4;
  p.fixup = (f32 && !(options & TG_OPT_NO_FP64_FIXUP) && !(options & TG_OPT_NO_CELL_SHIFT)) ? 1u
                                                                                            : 0u;
  p.tau = g->tau;
  p.mapping = tg::choose_mapping(g->dtype, b.tdim, g->gdim, b.k, p.nq, tile, options, p.mask);
  p.xcd_remap = tg::choose_xcd_remap();
  const bool relative = (options & TG_OPT_POINTS_RELATIVE) != 0;
  for (int i = 0; i < 3; ++i)
    p.origin[i] = static_cast<T>(i < g->gdim && !relative ? g->origin[i] : 0.0);
  return p;
}

tg::FactorOutputs outputs(const tg_outputs& o) {
  return tg::FactorOutputs{o.points, o.jacobians, o.detj, o.pinv, o.facet_normals, o.normals};
}

template <typename T>
int upload_persistent(tg_plan* plan) {
  const tg_geometry* g = plan->g;
  std::vector<tg::HybridBlock<T>> hb;
  std::vector<double> cost;  // per plan block: work per lane
  for (const PlanBlock& pb : plan->blocks) {
    const tg::Block& b = g->blocks[pb.index];
    tg::HybridBlock<T> h{};
    h.nodes = static_cast<const T*>(g->nodes);
    h.dofmap = b.dofmap;
    h.flags = b.flags;
    h.tab = static_cast<const T*>(pb.tabs->tab);
    h.tab64 = pb.tabs->tab64;
    h.ref_normals = static_cast<const T*>(pb.tabs->ref_normals);
    h.out = outputs(pb.out);
    h.p = params<T>(g, b, pb.pts, pb.mask, plan->tile, plan->options);
    h.shape = static_cast<uint32_t>(tg::hybrid_shape(b.tdim, g->gdim, b.k));
    hb.push_back(h);
    // Work of one item (one tile per wavefront, four points at a time): per group of four
    // points a fixed part (det J, staging, the output) and one part per node; fitted to the
    // kernel's times on single-block meshes (P2 tetrahedra, Q2 hexahedra, P2 prisms, P3
    // triangles: 38 node-equivalents of fixed work per point group).
    cost.push_back(static_cast<double>((pb.pts->nq + 3) / 4) * (38.0 + b.k));
  }
  // Work items (one tile of 16 cells per wavefront) block by block.
  std::vector<tg::WorkItem> items;
  std::vector<double> item_cost;
  for (size_t i = 0; i < hb.size(); ++i)
    for (uint64_t c = 0; c < hb[i].p.padded_cells; c += 16 * tg::kHybridWaves) {
      items.push_back(tg::WorkItem{static_cast<uint32_t>(i), static_cast<uint32_t>(c)});
      item_cost.push_back(cost[i]);
    }
  plan->num_items = static_cast<uint32_t>(items.size());
  size_t max_table = 0;
  uint32_t max_k = 1;
  for (const auto& h : hb) {
    max_table = std::max<size_t>(max_table, static_cast<size_t>(h.p.nq) * h.p.k * 4 * sizeof(T));
    max_k = std::max(max_k, h.p.k);
  }
  plan->xs_elems = tg::hybrid_xs_elems(max_k, g->gdim);
  plan->lds = tg::hybrid_lds_bytes(sizeof(T), plan->xs_elems, max_table);
  hipError_t e = tg::hybrid_grid<T>(plan->shift, plan->lds, &plan->grid);
  if (e == hipErrorInvalidValue) {  // too much LDS for one workgroup: launch block by block
    (void)hipGetLastError();
    plan->persistent = false;
    plan->num_items = 0;
    return TG_OK;
  }
  if (e != hipSuccess) return tg::hip_fail(e, "plan occupancy");
  plan->grid = std::max(1u, std::min(plan->grid, plan->num_items));
  // One contiguous range of about equal cost per workgroup, in item (block) order.
  double total = 0.0;
  for (double c : item_cost) total += c;
  const uint32_t G = plan->grid;
  std::vector<uint32_t> ranges(G + 1, static_cast<uint32_t>(items.size()));
  ranges[0] = 0;
  double acc = 0.0;
  uint32_t r = 1;
  for (size_t i = 0; i < items.size() && r < G; ++i) {
    acc += item_cost[i];
    while (r < G && acc >= total * r / G) ranges[r++] = static_cast<uint32_t>(i + 1);
  }
  e = hipMalloc(&plan->d_blocks, std::max<size_t>(1, hb.size()) * sizeof(hb[0]));
  if (e == hipSuccess)
    e = hipMalloc(reinterpret_cast<void**>(&plan->d_items),
                  std::max<size_t>(1, items.size()) * sizeof(tg::WorkItem));
  if (e == hipSuccess)
    e = hipMalloc(reinterpret_cast<void**>(&plan->d_ranges), (G + 1) * sizeof(uint32_t));
  if (e == hipSuccess && !hb.empty())
    e = hipMemcpy(plan->d_blocks, hb.data(), hb.size() * sizeof(hb[0]), hipMemcpyHostToDevice);
  if (e == hipSuccess && !items.empty())
    e = hipMemcpy(plan->d_items, items.data(), items.size() * sizeof(tg::WorkItem),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(plan->d_ranges, ranges.data(), (G + 1) * sizeof(uint32_t),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) return tg::hip_fail(e, "plan upload");
  return TG_OK;
}

// Output bytes of a block (the order in which a forked per-block plan launches).
uint64_t block_bytes(const tg_geometry* g, const PlanBlock& pb) {
  const tg::Block& b = g->blocks[pb.index];
  uint64_t len = 0;
  const uint32_t lens[6] = {static_cast<uint32_t>(g->gdim),
                            static_cast<uint32_t>(g->gdim * b.tdim), 1,
                            static_cast<uint32_t>(g->gdim * b.tdim), static_cast<uint32_t>(g->gdim),
                            static_cast<uint32_t>(g->gdim)};
  for (int i = 0; i < 6; ++i)
    if (pb.mask & (1u << i)) len += lens[i];
  return len * b.num_cells * pb.pts->nq;
}

template <typename T>
int launch_blocks(tg_plan* plan, hipStream_t s, bool fork) {
  const tg_geometry* g = plan->g;
  fork = fork && !plan->aux.empty();
  if (fork && hipEventRecord(plan->events[0], s) != hipSuccess)
    return tg::hip_fail(hipGetLastError(), "plan fork");
  for (size_t i = 0; i < plan->blocks.size(); ++i) {
    const PlanBlock& pb = plan->blocks[i];
    hipStream_t bs = s;
    if (fork && i > 0) {
      bs = plan->aux[i - 1];
      hipError_t e = hipStreamWaitEvent(bs, plan->events[0], 0);
      if (e != hipSuccess) return tg::hip_fail(e, "plan fork");
    }
    int rc = tg::enqueue_block<T>(g, g->blocks[pb.index], pb.pts, pb.tabs, pb.mask, plan->tile,
                                  plan->options, &pb.out, bs);
    if (rc) return rc;
    if (fork && i > 0) {
      hipError_t e = hipEventRecord(plan->events[i], bs);
      if (e == hipSuccess) e = hipStreamWaitEvent(s, plan->events[i], 0);
      if (e != hipSuccess) return tg::hip_fail(e, "plan join");
    }
  }
  return TG_OK;
}

// `fork`: per-block plans launch blocks 1.. on the auxiliary streams (used under graph capture,
// where the fork and join become graph edges; eagerly the extra event calls cost more host time
// than the overlap gains).
int plan_enqueue(tg_plan* plan, hipStream_t s, bool fork) {
  if (!plan->persistent)
    return plan->g->dtype == TG_F32 ? launch_blocks<float>(plan, s, fork)
                                    : launch_blocks<double>(plan, s, fork);
  if (plan->num_items == 0) return TG_OK;
  hipError_t e =
      plan->g->dtype == TG_F32
          ? tg::launch_hybrid<float>(plan->shift,
                                     static_cast<const tg::HybridBlock<float>*>(plan->d_blocks),
                                     plan->d_items, plan->d_ranges, plan->grid, plan->lds,
                                     plan->xs_elems, s)
          : tg::launch_hybrid<double>(plan->shift,
                                      static_cast<const tg::HybridBlock<double>*>(plan->d_blocks),
                                      plan->d_items, plan->d_ranges, plan->grid, plan->lds,
                                      plan->xs_elems, s);
  return e == hipSuccess ? TG_OK : tg::hip_fail(e, "persistent plan launch");
}

}  // namespace

extern "C" {

int tg_plan_create(const tg_geometry* g, tg_pointset* const* pts, uint32_t factors,
                   uint32_t tile, uint32_t options, const tg_outputs* outs, tg_plan** out) {
  if (!g || !pts || !outs || !out) return tg::fail(TG_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (tile != 1 && tile != 2 && tile != 4 && tile != 8 && tile != 16)
    return tg::fail(TG_ERR_INVALID_ARGUMENT, "tile must be 1, 2, 4, 8 or 16");
  const uint32_t all = TG_POINTS | TG_JACOBIAN | TG_DETJ | TG_PINV | TG_FACET_NORMALS | TG_NORMALS;
  if (factors == 0 || (factors & ~all)) return tg::fail(TG_ERR_INVALID_ARGUMENT, "bad factor mask");
  auto plan = std::make_unique<tg_plan>();
  plan->g = g;
  plan->tile = tile;
  plan->options = options;
  plan->shift = g->dtype == TG_F32 && !(options & TG_OPT_NO_CELL_SHIFT);
  // The persistent kernel reads the geometry's own node array: not with an FP64 copy attached.
  bool all_hybrid = !(options & TG_OPT_FORCE_GENERIC) && (g->dtype == TG_F64 || plan->shift) &&
                    !(plan->shift && g->nodes64);
  for (size_t bi = 0; bi < g->blocks.size(); ++bi) {
    const tg::Block& b = g->blocks[bi];
    tg_pointset* p = pts[bi];
    if (!p) return tg::fail(TG_ERR_CELL_NOT_IN_RULE, std::string("CellNotInRule(\"") +
                                                          tg::cell_name(b.cell) + "\")");
    if (p->cell != b.cell)
      return tg::fail(TG_ERR_CELL_NOT_IN_RULE,
                      std::string("CellNotInRule(\"") + tg::cell_name(b.cell) + "\")");
    uint32_t allowed = TG_POINTS | TG_JACOBIAN | TG_DETJ | TG_PINV;
    if (b.tdim == g->gdim && p->has_normals) allowed |= TG_FACET_NORMALS;
    if (b.tdim + 1 == g->gdim) allowed |= TG_NORMALS;
    const uint32_t mask = factors & allowed;
    if (mask == 0 || b.num_cells == 0 || p->nq == 0) continue;
    const void* ptrs[6] = {outs[bi].points, outs[bi].jacobians, outs[bi].detj,
                           outs[bi].pinv, outs[bi].facet_normals, outs[bi].normals};
    for (int i = 0; i < 6; ++i)
      if ((mask & (1u << i)) && !ptrs[i])
        return tg::fail(TG_ERR_INVALID_ARGUMENT,
                        "block " + std::to_string(bi) + ": missing output for a requested factor");
    PlanBlock pb{static_cast<int>(bi), p, nullptr, mask, outs[bi]};
    int rc = tg::device_tables(p, b.degree, b.family, g->dtype, &pb.tabs);
    if (rc) return rc;
    // The persistent kernel: a shape of its list, tiles of 16, the table within its LDS budget;
    // blocks that tg_eval_block runs on the dense matrix-core kernel keep to their own launch,
    // so that the plan's fields equal it bit for bit.
    if (tg::hybrid_shape(b.tdim, g->gdim, b.k) < 0 || tile != 16 ||
        p->nq * b.k * 4 * (g->dtype == TG_F32 ? 4 : 8) > tg::kMaxLdsTableBytes ||
        tg::takes_dense_mfma(g, b, pb.tabs, mask, tile, options))
      all_hybrid = false;
    plan->blocks.push_back(pb);
  }
  plan->persistent =
      all_hybrid && (options & TG_OPT_PLAN_PERSISTENT) && !(options & TG_OPT_PLAN_PER_BLOCK);
  if (!plan->persistent && plan->blocks.size() > 1 && !(options & TG_OPT_PLAN_SERIAL)) {
    // Largest block first, on the capturing stream; the others on their own streams.
    std::stable_sort(plan->blocks.begin(), plan->blocks.end(),
                     [&](const PlanBlock& a, const PlanBlock& b) {
                       return block_bytes(g, a) > block_bytes(g, b);
                     });
    plan->events.resize(plan->blocks.size(), nullptr);
    for (hipEvent_t& ev : plan->events) {
      hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
      if (e != hipSuccess) return tg::hip_fail(e, "plan events");
    }
    plan->aux.resize(plan->blocks.size() - 1, nullptr);
    for (hipStream_t& a : plan->aux) {
      hipError_t e = hipStreamCreateWithFlags(&a, hipStreamNonBlocking);
      if (e != hipSuccess) return tg::hip_fail(e, "plan streams");
    }
  }
  if (plan->persistent) {
    int rc = g->dtype == TG_F32 ? upload_persistent<float>(plan.get())
                                : upload_persistent<double>(plan.get());
    if (rc) return rc;
  }
  *out = plan.release();
  return TG_OK;
}

int tg_plan_destroy(tg_plan* plan) {
  delete plan;
  return TG_OK;
}

int tg_plan_info(const tg_plan* plan, uint32_t* num_items, uint32_t* grid, int32_t* persistent) {
  if (!plan) return tg::fail(TG_ERR_INVALID_ARGUMENT, "null plan");
  if (num_items) *num_items = plan->num_items;
  if (grid) *grid = plan->grid;
  if (persistent) *persistent = plan->persistent ? 1 : 0;
  return TG_OK;
}

int tg_plan_launch(tg_plan* plan, void* stream) {
  if (!plan) return tg::fail(TG_ERR_INVALID_ARGUMENT, "null plan");
  return plan_enqueue(plan, static_cast<hipStream_t>(stream), false);
}

int tg_plan_graph_launch(tg_plan* plan, void* stream) {
  if (!plan) return tg::fail(TG_ERR_INVALID_ARGUMENT, "null plan");
  hipError_t e = hipSuccess;
  if (!plan->exec) {
    // Capture on a private stream of default priority, once; replays launch on `stream`.
    if (!plan->capture) e = hipStreamCreateWithFlags(&plan->capture, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(plan->capture, hipStreamCaptureModeThreadLocal);
    int rc = TG_OK;
    if (e == hipSuccess) {
      rc = plan_enqueue(plan, plan->capture, true);
      // End the capture whatever the enqueue returned, so the stream leaves capture mode.
      hipError_t end = hipStreamEndCapture(plan->capture, &plan->graph);
      if (rc == TG_OK) e = end;
    }
    if (rc == TG_OK && e == hipSuccess)
      e = hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0);
    if (rc != TG_OK || e != hipSuccess) {
      // Release the partial graph (the capture stream is kept for the next attempt).
      if (plan->graph) hipGraphDestroy(plan->graph);
      plan->graph = nullptr;
      plan->exec = nullptr;
      return rc != TG_OK ? rc : tg::hip_fail(e, "plan graph capture");
    }
  }
  e = hipGraphLaunch(plan->exec, static_cast<hipStream_t>(stream));
  return e == hipSuccess ? TG_OK : tg::hip_fail(e, "plan graph launch");
}

}  // extern "C"
