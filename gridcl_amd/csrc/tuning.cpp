// The launch mapping of the register-resident factor kernels per (precision, tdim, gdim, nodes,
// points, output class): the table written by the autotuner (gridcl_amd/tune.py), compiled in as the
// reference compiles in its measured degree policy (tetraloma-geometry/degree-policy.toml,
// strategy.rs:373-407), with the same two overrides: an option bit of the call, then the
// environment variable TETRALOMA_GEOM_MAPPING (cell | cell_staged | tile).
#include <cstdlib>
#include <cstring>

#include "../../include/tetraloma_geom.h"
#include "geom_internal.hpp"

namespace tg {
namespace {

// Output classes: 0 for calls writing more than three components per point (X, J and det J
// together, J or K alone), 1 for at most three (det J, X, normals or facet normals alone), whose
// stores are few enough that one lane per cell measures faster than a tile per wavefront.
struct TuneEntry {
  int dtype, tdim, gdim;
  uint32_t k, nq_lo, nq_hi;  // points per cell in [nq_lo, nq_hi]
  uint32_t cls;              // output class
  uint32_t mapping;
};

#include "tuning_table.inc"

uint32_t env_mapping() {
  static const uint32_t v = [] {
    const char* e = std::getenv("TETRALOMA_GEOM_MAPPING");
    if (!e) return 0xffffffffu;
    if (!std::strcmp(e, "cell")) return static_cast<uint32_t>(kMapCell);
    if (!std::strcmp(e, "cell_staged")) return static_cast<uint32_t>(kMapCellStaged);
    if (!std::strcmp(e, "tile")) return static_cast<uint32_t>(kMapTile);
    return 0xffffffffu;
  }();
  return v;
}

}  // namespace

uint32_t choose_xcd_remap() {
  static const uint32_t v = [] {
    const char* e = std::getenv("TETRALOMA_GEOM_XCD_REMAP");
    return (e && e[0] == '0') ? 0u : 1u;
  }();
  return v;
}

uint32_t output_components(uint32_t factors, int tdim, int gdim) {
  uint32_t n = 0;
  if (factors & TG_POINTS) n += gdim;
  if (factors & TG_JACOBIAN) n += tdim * gdim;
  if (factors & TG_DETJ) n += 1;
  if (factors & TG_PINV) n += tdim * gdim;
  if (factors & TG_FACET_NORMALS) n += gdim;
  if (factors & TG_NORMALS) n += gdim;
  return n;
}

uint32_t choose_mapping(int dtype, int tdim, int gdim, uint32_t k, uint32_t nq, uint32_t tile,
                        uint32_t options, uint32_t factors) {
  if (tile != 16) return kMapCell;
  if (options & TG_OPT_MAP_CELL) return kMapCell;
  if (options & TG_OPT_MAP_CELL_STAGED) return kMapCellStaged;
  if (options & TG_OPT_MAP_TILE) {
    const size_t bytes = static_cast<size_t>(nq) * k * 4 * (dtype == TG_F64 ? 8 : 4);
    return bytes > kMaxLdsTableBytes ? static_cast<uint32_t>(kMapCellStaged)
                                     : static_cast<uint32_t>(kMapTile);
  }
  const uint32_t env = env_mapping();
  if (env != 0xffffffffu) {
    const size_t bytes = static_cast<size_t>(nq) * k * 4 * (dtype == TG_F64 ? 8 : 4);
    return (env == kMapTile && bytes > kMaxLdsTableBytes) ? static_cast<uint32_t>(kMapCellStaged)
                                                          : env;
  }
  // The entry of the call's output class, else that of class 0 (the fused factors).
  const uint32_t cls = output_components(factors, tdim, gdim) <= 3 ? 1u : 0u;
  uint32_t m = kDefaultMapping;
  bool found = false;
  for (uint32_t c = cls; !found; c = 0) {
    for (const TuneEntry& e : kTuned)
      if (e.dtype == dtype && e.tdim == tdim && e.gdim == gdim && e.k == k && nq >= e.nq_lo &&
          nq <= e.nq_hi && e.cls == c) {
        m = e.mapping;
        found = true;
        break;
      }
    if (c == 0) break;
  }
  const size_t table_bytes = static_cast<size_t>(nq) * k * 4 * (dtype == TG_F64 ? 8 : 4);
  if (m == kMapTile && table_bytes > kMaxLdsTableBytes) m = kMapCellStaged;
  return m;
}

}  // namespace tg
