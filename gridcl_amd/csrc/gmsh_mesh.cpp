// The mesh a Gmsh file describes (tetraloma-io/src/gmsh/convert.rs:169-356, to_gmsh_mesh, with
// the layout, serendipity completion and dimension inference of chunk.rs:77-578, 659-711), and
// the C-ABI entry points of the reader.
//
// Rules, in the reference's order of checks:
// - node tags, then element tags, are unique over the file (DuplicateOriginId, the first repeat
//   in file order);
// - the topological dimension is the highest dimension of a kept, non-empty element block; its
//   elements form one cell block per cell type in cell-type order (one degree per type,
//   DuplicateCellBlock otherwise), cells in file order, nodes permuted from Gmsh's order to A.5;
// - the geometric dimension is 3 if any |z| exceeds 1e-12 times the bounding-box diameter, else 2
//   if any |y| does, else 1, at least the topological dimension, or the caller's if not below;
// - serendipity cells become complete P2: a missing face-centre node is shared per face (keyed by
//   its sorted corner tags, reusing a complete-P2 neighbour's node) and otherwise added with the
//   next tag above the file's largest, in order of first use;
// - lower-dimensional elements become tag groups (corner tags in internal vertex order); each
//   corner must be a node (Parse) and a vertex of some cell (UnmatchedTagElement);
// - elementary tags are entity tags (a partitioned entity reports its parent's), physical tags
//   come from the entity; elements of partitioned entities whose parent has a higher dimension
//   (partition interfaces) are dropped.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <map>
#include <memory>
#include <unordered_map>
#include <unordered_set>

#include "../../include/tetraloma_geom.h"
#include "gmsh_internal.hpp"

namespace tg::gmsh {

namespace {

constexpr int kParse = TG_ERR_PARSE;

Error content_error(const std::string& m) {
  return Error{kParse, "Parse: Gmsh: " + m + " (no byte offset: the file's content)", 0};
}

void check_len(size_t found, size_t count, size_t per) {
  if (per != 0 && count > SIZE_MAX / per) throw Error{TG_ERR_TOO_LARGE, "TooLarge", 0};
  const size_t expected = count * per;
  if (found != expected)
    throw Error{TG_ERR_LENGTH_MISMATCH,
                "LengthMismatch: expected " + std::to_string(expected) + ", found " +
                    std::to_string(found),
                static_cast<int64_t>(found)};
}

// DuplicateOriginId at the first repeated tag in order: a bitmap over the tags' range when they
// are compact (the usual 1..n numbering), a hash set otherwise.
void check_unique(const std::vector<int64_t>& tags) {
  if (tags.empty()) return;
  const auto mm = std::minmax_element(tags.begin(), tags.end());
  const int64_t lo = *mm.first;
  const uint64_t span = static_cast<uint64_t>(*mm.second - lo);
  auto duplicate = [](int64_t t) {
    return Error{TG_ERR_DUPLICATE_ORIGIN_ID, "DuplicateOriginId: duplicate origin id " +
                                                 std::to_string(t), t};
  };
  if (span < 8 * tags.size() + 4096) {
    std::vector<uint64_t> seen(span / 64 + 1, 0);
    for (int64_t t : tags) {
      const uint64_t i = static_cast<uint64_t>(t - lo);
      const uint64_t bit = 1ull << (i & 63);
      if (seen[i >> 6] & bit) throw duplicate(t);
      seen[i >> 6] |= bit;
    }
    return;
  }
  std::unordered_set<int64_t> seen;
  seen.reserve(tags.size());
  for (int64_t t : tags)
    if (!seen.insert(t).second) throw duplicate(t);
}

Error missing_node(int64_t t, int64_t element) {
  return content_error("element " + std::to_string(element) + ": node " + std::to_string(t) +
                       " is not in $Nodes");
}

struct Info {
  int32_t elementary;
  std::vector<int32_t> physical;  // sorted, distinct
  bool keep;
};

std::vector<int32_t> sorted(std::vector<int32_t> v) {
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  return v;
}

class Layout {
 public:
  int tdim = 0;
  std::vector<std::pair<Cell, unsigned>> blocks;  // cell-type order

  explicit Layout(const File& f) {
    if (f.has_entities)
      for (int dim = 0; dim < 4; ++dim)
        for (const Entity& e : f.entities[dim])
          info_[{dim, e.tag}] = Info{e.tag, sorted(e.physical), true};
    if (f.has_partitioned)
      for (int dim = 0; dim < 4; ++dim)
        for (const PartitionedEntity& p : f.partitioned[dim])
          info_[{dim, p.entity.tag}] =
              Info{p.parent_tag, sorted(p.entity.physical), p.parent_dim == dim};
    std::vector<std::pair<const ElementBlock*, ElementType>> kept;
    for (const ElementBlock& b : f.element_blocks) {
      const ElementType ty = element_type(b.element_type);
      if (tg::tdim(ty.cell) != b.dim)
        throw content_error("element type " + std::to_string(b.element_type) +
                            " in an entity of dimension " + std::to_string(b.dim));
      if (keeps(b)) kept.push_back({&b, ty});
    }
    for (const auto& k : kept)
      if (!k.first->tags.empty()) tdim = std::max(tdim, k.first->dim);
    if (tdim == 0) throw content_error("no elements of dimension 1 or more");
    for (const auto& k : kept) {
      if (k.first->dim != tdim) continue;
      auto it = std::find_if(blocks.begin(), blocks.end(),
                             [&](const auto& b) { return b.first == k.second.cell; });
      if (it == blocks.end())
        blocks.push_back({k.second.cell, k.second.degree});
      else if (it->second != k.second.degree)
        throw Error{TG_ERR_DUPLICATE_CELL_BLOCK,
                    std::string("DuplicateCellBlock: ") + cell_name(k.second.cell), 0};
    }
    std::sort(blocks.begin(), blocks.end());
  }

  // The elementary tag, physical tags and kept flag of the elements of `b`; entities missing
  // from $Entities have their own tag and no physical tags.
  Info entity(const ElementBlock& b) const {
    auto it = info_.find({b.dim, b.entity_tag});
    if (it == info_.end()) return Info{b.entity_tag, {}, true};
    return it->second;
  }
  bool keeps(const ElementBlock& b) const {
    auto it = info_.find({b.dim, b.entity_tag});
    return it == info_.end() || it->second.keep;
  }
  bool top(const ElementBlock& b) const { return b.dim == tdim && keeps(b); }

 private:
  std::map<std::pair<int, int32_t>, Info> info_;
};

// Node tag -> row: a dense table when the tags are compact, a hash map otherwise.
class NodeIndex {
 public:
  void build(const std::vector<int64_t>& tags) {
    lo_ = INT64_MAX;
    int64_t hi = INT64_MIN;
    for (int64_t t : tags) {
      lo_ = std::min(lo_, t);
      hi = std::max(hi, t);
    }
    dense_.clear();
    map_.clear();
    if (!tags.empty() && static_cast<uint64_t>(hi - lo_) < 4 * tags.size() + 1024) {
      dense_.assign(static_cast<size_t>(hi - lo_) + 1, UINT32_MAX);
      for (size_t i = 0; i < tags.size(); ++i) dense_[tags[i] - lo_] = static_cast<uint32_t>(i);
      use_dense_ = true;
    } else {
      map_.reserve(tags.size());
      for (size_t i = 0; i < tags.size(); ++i) map_.emplace(tags[i], static_cast<uint32_t>(i));
      use_dense_ = false;
    }
  }
  // The row of tag t, or UINT32_MAX.
  uint32_t find(int64_t t) const {
    if (use_dense_) {
      if (t < lo_ || static_cast<uint64_t>(t - lo_) >= dense_.size()) return UINT32_MAX;
      return dense_[t - lo_];
    }
    auto it = map_.find(t);
    return it == map_.end() ? UINT32_MAX : it->second;
  }
  uint32_t get(int64_t t, int64_t element) const {
    const uint32_t r = find(t);
    if (r == UINT32_MAX) throw missing_node(t, element);
    return r;
  }

 private:
  bool use_dense_ = true;
  int64_t lo_ = 0;
  std::vector<uint32_t> dense_;
  std::unordered_map<int64_t, uint32_t> map_;
};

// D_GEO of the nodes: 3 if any |z| > 1e-12 L, else 2 if any |y| > 1e-12 L, else 1, with L the
// bounding-box diameter (chunk.rs:659-693).
int infer_gdim(const File& f) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  double mx[3] = {0, 0, 0};
  for (const NodeBlock& b : f.node_blocks)
    for (size_t i = 0; i < b.tags.size(); ++i)
      for (int g = 0; g < 3; ++g) {
        const double x = b.coords[3 * i + g];
        lo[g] = std::fmin(lo[g], x);
        hi[g] = std::fmax(hi[g], x);
        mx[g] = std::fmax(mx[g], std::fabs(x));
      }
  double d2 = 0.0;
  for (int g = 0; g < 3; ++g)
    if (hi[g] >= lo[g]) d2 += (hi[g] - lo[g]) * (hi[g] - lo[g]);
  const double threshold = 1e-12 * std::sqrt(d2);
  if (mx[2] > threshold) return 3;
  if (mx[1] > threshold) return 2;
  return 1;
}

// The number of P2 Lagrange nodes in one sub-entity's interior (A.5 blocks at degree 2).
int p2_block(Cell t) {
  return (t == kPoint || t == kInterval || t == kQuadrilateral || t == kHexahedron) ? 1 : 0;
}

// (dim, index) of the sub-entity whose interior holds internal node `node` of P2 `cell`.
std::pair<int, int> entity_of_node(Cell cell, int node) {
  int start = 0;
  for (int dim = 0; dim <= tg::tdim(cell); ++dim)
    for (int e = 0; e < num_sub_entities(cell, dim); ++e) {
      const int n = p2_block(sub_entity_type(cell, dim, e));
      if (node >= start && node < start + n) return {dim, e};
      start += n;
    }
  return {-1, -1};
}

// The first A.5 index of face f's interior at P2, or -1 without one.
int p2_face_interior(Cell cell, int f) {
  int start = 0;
  for (int dim = 0; dim <= 2; ++dim)
    for (int e = 0; e < num_sub_entities(cell, dim); ++e) {
      const int n = p2_block(sub_entity_type(cell, dim, e));
      if (dim == 2 && e == f) return n > 0 ? start : -1;
      start += n;
    }
  return -1;
}

std::vector<int64_t> face_key(Cell cell, int f, const std::vector<int64_t>& internal) {
  std::vector<int64_t> key;
  for (int v : sub_entity(cell, 2, f)) key.push_back(internal[v]);
  std::sort(key.begin(), key.end());
  return key;
}

struct KeyHash {
  size_t operator()(const std::vector<int64_t>& k) const {
    size_t h = 1469598103934665603ull;
    for (int64_t v : k) h = (h ^ static_cast<size_t>(v)) * 1099511628211ull;
    return h;
  }
};

// Gmsh-to-A.5 map of (cell, degree) with k entries, or UnsupportedDegree.
const std::vector<int>& node_map(Cell cell, unsigned degree, size_t k) {
  const std::vector<int>& m = gmsh_to_internal(cell, degree);
  if (m.size() != k)
    throw Error{TG_ERR_UNSUPPORTED_DEGREE,
                std::string("UnsupportedDegree: no node order for ") + cell_name(cell) +
                    " of degree " + std::to_string(degree),
                degree};
  return m;
}

// The file with its serendipity cells completed to P2 (chunk.rs:504-578); false if it has none.
bool complete(const File& file, const Layout& layout, File& out) {
  bool any = false;
  for (const ElementBlock& b : file.element_blocks) {
    if (!layout.top(b)) continue;
    const ElementType ty = element_type(b.element_type);
    check_len(b.nodes.size(), b.tags.size(), ty.num_nodes);
    any = any || ty.serendipity;
  }
  if (!any) return false;
  std::vector<int64_t> tags;
  std::vector<double> coords;
  int64_t next_tag = 1;
  for (const NodeBlock& b : file.node_blocks)
    for (size_t i = 0; i < b.tags.size(); ++i) {
      tags.push_back(b.tags[i]);
      coords.insert(coords.end(), &b.coords[3 * i], &b.coords[3 * i + 3]);
      if (b.tags[i] == INT64_MAX) throw Error{TG_ERR_TOO_LARGE, "TooLarge", 0};
      next_tag = std::max(next_tag, b.tags[i] + 1);
    }
  const size_t num_file_nodes = tags.size();
  NodeIndex index;
  index.build(tags);
  std::unordered_map<int64_t, uint32_t> added;  // tags of added nodes -> rows

  std::unordered_map<std::vector<int64_t>, int64_t, KeyHash> faces;
  for (const auto& blk : layout.blocks) {
    const Cell cell = blk.first;
    if (blk.second != 2 || tg::tdim(cell) != 3) continue;
    const std::vector<int>& map = node_map(cell, 2, num_lagrange_nodes(cell, 2));
    for (const ElementBlock& b : file.element_blocks) {
      if (!layout.top(b)) continue;
      const ElementType ty = element_type(b.element_type);
      if (ty.cell != cell || ty.serendipity) continue;
      std::vector<int64_t> internal(ty.num_nodes);
      for (size_t r = 0; r < b.tags.size(); ++r) {
        for (size_t k = 0; k < ty.num_nodes; ++k) internal[map[k]] = b.nodes[r * ty.num_nodes + k];
        for (int f = 0; f < num_sub_entities(cell, 2); ++f) {
          const int interior = p2_face_interior(cell, f);
          if (interior >= 0) faces.emplace(face_key(cell, f, internal), internal[interior]);
        }
      }
    }
  }

  out = file;
  for (const auto& blk : layout.blocks) {
    const Cell cell = blk.first;
    const ElementType full = lagrange_element_type(cell, 2);
    const std::vector<int>& map = node_map(cell, 2, full.num_nodes);
    size_t nf = 0, nn = 0;
    const std::vector<double>& w = serendipity_weights(cell, &nf, &nn);
    for (size_t bi = 0; bi < file.element_blocks.size(); ++bi) {
      const ElementBlock& b = file.element_blocks[bi];
      const ElementType ty = element_type(b.element_type);
      if (!layout.top(b) || ty.cell != cell || !ty.serendipity) continue;
      std::vector<int64_t> rewritten;
      rewritten.reserve(b.tags.size() * full.num_nodes);
      std::vector<int64_t> internal(nn);
      std::vector<std::array<double, 3>> x(nf);
      for (size_t r = 0; r < b.tags.size(); ++r) {
        const int64_t element = b.tags[r];
        const int64_t* row = &b.nodes[r * ty.num_nodes];
        for (size_t k = 0; k < nf; ++k) internal[map[k]] = row[k];
        for (size_t k = 0; k < nf; ++k) {
          uint32_t i = index.find(row[k]);
          if (i == UINT32_MAX) {
            auto a = added.find(row[k]);
            if (a == added.end()) throw missing_node(row[k], element);
            i = a->second;
          }
          for (int g = 0; g < 3; ++g) x[k][g] = coords[3 * static_cast<size_t>(i) + g];
        }
        for (size_t m = 0; m < nn - nf; ++m) {
          const int node = map[nf + m];
          const std::pair<int, int> ent = entity_of_node(cell, node);
          const bool keyed = ent.first == 2 && tg::tdim(cell) == 3;
          std::vector<int64_t> key;
          if (keyed) {
            key = face_key(cell, ent.second, internal);
            auto hit = faces.find(key);
            if (hit != faces.end()) {
              internal[node] = hit->second;
              continue;
            }
          }
          double p[3] = {0.0, 0.0, 0.0};
          for (size_t j = 0; j < nf; ++j)
            for (int g = 0; g < 3; ++g) p[g] += w[m * nf + j] * x[j][g];
          const int64_t t = next_tag;
          if (t == INT64_MAX) throw Error{TG_ERR_TOO_LARGE, "TooLarge", 0};
          next_tag = t + 1;
          added.emplace(t, static_cast<uint32_t>(tags.size()));
          tags.push_back(t);
          coords.insert(coords.end(), p, p + 3);
          internal[node] = t;
          if (keyed) faces.emplace(std::move(key), t);
        }
        for (int i : map) rewritten.push_back(internal[i]);
      }
      out.element_blocks[bi].element_type = full.id;
      out.element_blocks[bi].nodes = std::move(rewritten);
    }
  }
  if (tags.size() > num_file_nodes) {
    NodeBlock nb;
    nb.dim = layout.tdim;
    nb.entity_tag = 0;
    for (const ElementBlock& b : file.element_blocks)
      if (layout.top(b)) {
        nb.entity_tag = b.entity_tag;
        break;
      }
    nb.tags.assign(tags.begin() + num_file_nodes, tags.end());
    nb.coords.assign(coords.begin() + 3 * num_file_nodes, coords.end());
    out.node_blocks.push_back(std::move(nb));
  }
  return true;
}

}  // namespace

Mesh to_mesh(const File& file_in, int requested_gdim) {
  {
    std::vector<int64_t> all;
    for (const NodeBlock& b : file_in.node_blocks) all.insert(all.end(), b.tags.begin(), b.tags.end());
    check_unique(all);
    all.clear();
    for (const ElementBlock& b : file_in.element_blocks)
      all.insert(all.end(), b.tags.begin(), b.tags.end());
    check_unique(all);
  }
  for (const ElementBlock& b : file_in.element_blocks) {
    const ElementType ty = element_type(b.element_type);
    if (tg::tdim(ty.cell) != b.dim)
      throw content_error("element type " + std::to_string(b.element_type) +
                          " in an entity of dimension " + std::to_string(b.dim));
    check_len(b.nodes.size(), b.tags.size(), ty.num_nodes);
  }
  for (const NodeBlock& b : file_in.node_blocks) check_len(b.coords.size(), 3 * b.tags.size(), 1);
  const Layout layout(file_in);
  Mesh mesh;
  mesh.tdim = layout.tdim;
  const int inferred = std::max(infer_gdim(file_in), mesh.tdim);
  if (requested_gdim == 0) {
    mesh.gdim = inferred;
  } else if (requested_gdim >= 1 && requested_gdim <= 3 && requested_gdim >= inferred) {
    mesh.gdim = requested_gdim;
  } else {
    throw Error{TG_ERR_DIMENSION_MISMATCH,
                "DimensionMismatch: expected (tdim " + std::to_string(mesh.tdim) + ", gdim " +
                    std::to_string(inferred) + "), found gdim " + std::to_string(requested_gdim),
                requested_gdim};
  }
  File completed;
  const File& file = complete(file_in, layout, completed) ? completed : file_in;

  size_t num_nodes = 0;
  for (const NodeBlock& b : file.node_blocks) num_nodes += b.tags.size();
  if (num_nodes > UINT32_MAX) throw Error{TG_ERR_TOO_LARGE, "TooLarge: too many nodes", 0};
  mesh.node_tags.reserve(num_nodes);
  mesh.nodes.reserve(num_nodes * mesh.gdim);
  for (const NodeBlock& b : file.node_blocks)
    for (size_t i = 0; i < b.tags.size(); ++i) {
      mesh.node_tags.push_back(b.tags[i]);
      for (int g = 0; g < mesh.gdim; ++g) mesh.nodes.push_back(b.coords[3 * i + g]);
    }
  NodeIndex index;
  index.build(mesh.node_tags);

  // Cell blocks.
  std::vector<char> is_vertex(num_nodes, 0);
  for (const auto& lb : layout.blocks) {
    CellBlock cb;
    cb.cell = lb.first;
    cb.degree = lb.second;
    const size_t nv = static_cast<size_t>(num_vertices(cb.cell));
    size_t reserve = 0;
    for (const ElementBlock& b : file.element_blocks)
      if (b.dim == layout.tdim && layout.keeps(b) && element_type(b.element_type).cell == cb.cell)
        reserve += b.tags.size();
    cb.nodes.reserve(reserve * num_lagrange_nodes(cb.cell, cb.degree));
    cb.tags.reserve(reserve);
    cb.elementary.reserve(reserve);
    for (const ElementBlock& b : file.element_blocks) {
      if (b.dim != layout.tdim || !layout.keeps(b)) continue;
      const ElementType ty = element_type(b.element_type);
      if (ty.cell != cb.cell) continue;
      check_len(b.nodes.size(), b.tags.size(), ty.num_nodes);
      if (ty.serendipity)
        throw content_error("element type " + std::to_string(ty.id) +
                            ": serendipity cells must be completed first");
      if (ty.degree != cb.degree)
        throw Error{TG_ERR_DUPLICATE_CELL_BLOCK,
                    std::string("DuplicateCellBlock: ") + cell_name(ty.cell), 0};
      const Info info = layout.entity(b);
      const std::vector<int>& map = node_map(ty.cell, ty.degree, ty.num_nodes);
      const size_t k = ty.num_nodes;
      std::vector<int64_t> row(k);
      for (size_t r = 0; r < b.tags.size(); ++r) {
        if (cb.tags.size() >= UINT32_MAX) throw Error{TG_ERR_TOO_LARGE, "TooLarge", 0};
        const int64_t c = static_cast<int64_t>(cb.tags.size());
        for (size_t j = 0; j < k; ++j) row[map[j]] = b.nodes[r * k + j];
        for (size_t j = 0; j < k; ++j) {
          const uint32_t i = index.get(row[j], b.tags[r]);
          cb.nodes.push_back(i);
          if (j < nv) is_vertex[i] = 1;
        }
        cb.tags.push_back(b.tags[r]);
        cb.elementary.push_back(info.elementary);
        for (int32_t p : info.physical) cb.physical.push_back({c, p});
      }
    }
    mesh.blocks.push_back(std::move(cb));
  }

  // Tag groups: every corner is a node and a cell vertex, checked in file order.
  for (const ElementBlock& b : file.element_blocks) {
    if (b.dim >= layout.tdim || !layout.keeps(b)) continue;
    const ElementType ty = element_type(b.element_type);
    const size_t nv = static_cast<size_t>(ty.cell == kPoint ? 1 : num_vertices(ty.cell));
    const std::vector<int> vmap =
        ty.cell == kPoint ? std::vector<int>{0} : node_map(ty.cell, 1, nv);
    for (size_t r = 0; r < b.tags.size(); ++r) {
      std::vector<int64_t> corners(nv);
      for (size_t k = 0; k < nv; ++k) corners[vmap[k]] = b.nodes[r * ty.num_nodes + k];
      for (int64_t t : corners) {
        const uint32_t i = index.get(t, b.tags[r]);
        if (!is_vertex[i])
          throw Error{TG_ERR_UNMATCHED_TAG_ELEMENT,
                      "UnmatchedTagElement: element " + std::to_string(b.tags[r]) +
                          " of dimension " + std::to_string(b.dim) +
                          " has a corner that is not a cell vertex",
                      b.tags[r]};
      }
    }
  }
  for (const ElementBlock& b : file.element_blocks) {
    if (b.dim >= layout.tdim || !layout.keeps(b)) continue;
    const Info info = layout.entity(b);
    const ElementType ty = element_type(b.element_type);
    check_len(b.nodes.size(), b.tags.size(), ty.num_nodes);
    const size_t nv = static_cast<size_t>(ty.cell == kPoint ? 1 : num_vertices(ty.cell));
    const std::vector<int> vmap =
        ty.cell == kPoint ? std::vector<int>{0} : node_map(ty.cell, 1, nv);
    auto it = std::find_if(mesh.tag_groups.begin(), mesh.tag_groups.end(),
                           [&](const TagGroup& g) { return g.cell == ty.cell; });
    if (it == mesh.tag_groups.end()) {
      TagGroup g;
      g.dim = b.dim;
      g.cell = ty.cell;
      mesh.tag_groups.push_back(std::move(g));
      it = mesh.tag_groups.end() - 1;
    }
    TagGroup& g = *it;
    std::vector<int64_t> corners(nv);
    for (size_t r = 0; r < b.tags.size(); ++r) {
      if (g.tags.size() >= UINT32_MAX) throw Error{TG_ERR_TOO_LARGE, "TooLarge", 0};
      const int64_t e = static_cast<int64_t>(g.tags.size());
      for (size_t k = 0; k < nv; ++k) corners[vmap[k]] = b.nodes[r * ty.num_nodes + k];
      g.corners.insert(g.corners.end(), corners.begin(), corners.end());
      g.tags.push_back(b.tags[r]);
      g.elementary.push_back(info.elementary);
      for (int32_t p : info.physical) g.physical.push_back({e, p});
    }
  }
  std::stable_sort(mesh.tag_groups.begin(), mesh.tag_groups.end(),
                   [](const TagGroup& a, const TagGroup& b) { return a.cell < b.cell; });
  for (TagGroup& g : mesh.tag_groups) std::sort(g.physical.begin(), g.physical.end());
  mesh.names = file.names;
  return mesh;
}

}  // namespace tg::gmsh

// ---- C ABI ------------------------------------------------------------------------------------

struct tg_gmsh_mesh {
  tg::gmsh::Mesh mesh;
  // What the file held as stored: its encoding, node count, element blocks and small sections.
  bool binary = false;
  uint64_t file_nodes = 0;
  std::vector<tg_gmsh_element_block> element_blocks;
  bool partitioned = false;
  uint64_t num_partitions = 0;
  uint64_t num_periodic = 0;
};

namespace {

thread_local int64_t g_detail = 0;

int read_into(const unsigned char* bytes, size_t len, int32_t gdim, tg_gmsh_mesh** out) {
  try {
    tg::gmsh::File f = tg::gmsh::parse(bytes, len);
    auto m = std::make_unique<tg_gmsh_mesh>();
    m->binary = f.binary;
    for (const auto& b : f.node_blocks) m->file_nodes += b.tags.size();
    for (const auto& b : f.element_blocks)
      m->element_blocks.push_back({b.dim, b.entity_tag, b.element_type, b.tags.size()});
    m->partitioned = f.has_partitioned;
    m->num_partitions = f.num_partitions;
    m->num_periodic = f.periodic.size();
    m->mesh = tg::gmsh::to_mesh(f, gdim);
    *out = m.release();
    g_detail = 0;
    return TG_OK;
  } catch (const tg::gmsh::Error& e) {
    g_detail = e.detail;
    return tg::fail(e.code, e.message);
  } catch (const std::bad_alloc&) {
    return tg::fail(TG_ERR_TOO_LARGE, "TooLarge: out of host memory reading the Gmsh file");
  }
}

}  // namespace

extern "C" {

int tg_gmsh_read_bytes(const void* bytes, uint64_t len, int32_t gdim, tg_gmsh_mesh** out) {
  if (!out || (!bytes && len)) return tg::fail(TG_ERR_INVALID_ARGUMENT, "null bytes or output");
  *out = nullptr;
  static const unsigned char empty = 0;
  return read_into(bytes ? static_cast<const unsigned char*>(bytes) : &empty,
                   static_cast<size_t>(len), gdim, out);
}

int tg_gmsh_read_file(const char* path, int32_t gdim, tg_gmsh_mesh** out) {
  if (!path || !out) return tg::fail(TG_ERR_INVALID_ARGUMENT, "null path or output");
  *out = nullptr;
  FILE* fp = std::fopen(path, "rb");
  if (!fp) return tg::fail(TG_ERR_IO, std::string("Io: ") + path + ": cannot open");
  std::vector<unsigned char> bytes;
  unsigned char buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, fp)) > 0) bytes.insert(bytes.end(), buf, buf + n);
  const bool bad = std::ferror(fp) != 0;
  std::fclose(fp);
  if (bad) return tg::fail(TG_ERR_IO, std::string("Io: ") + path + ": read error");
  static const unsigned char empty = 0;
  return read_into(bytes.empty() ? &empty : bytes.data(), bytes.size(), gdim, out);
}

int tg_gmsh_destroy(tg_gmsh_mesh* m) {
  delete m;
  return TG_OK;
}

int64_t tg_gmsh_last_error_detail(void) { return g_detail; }

int tg_gmsh_get_info(const tg_gmsh_mesh* m, tg_gmsh_info* info) {
  if (!m || !info) return tg::fail(TG_ERR_INVALID_ARGUMENT, "null mesh or output");
  info->tdim = m->mesh.tdim;
  info->gdim = m->mesh.gdim;
  info->num_nodes = m->mesh.node_tags.size();
  info->num_blocks = static_cast<int32_t>(m->mesh.blocks.size());
  info->num_tag_groups = static_cast<int32_t>(m->mesh.tag_groups.size());
  info->num_names = static_cast<int32_t>(m->mesh.names.size());
  info->binary = m->binary ? 1 : 0;
  info->file_nodes = m->file_nodes;
  info->num_element_blocks = static_cast<int32_t>(m->element_blocks.size());
  info->partitioned = m->partitioned ? 1 : 0;
  info->num_partitions = m->num_partitions;
  info->num_periodic = m->num_periodic;
  return TG_OK;
}

const double* tg_gmsh_nodes(const tg_gmsh_mesh* m) { return m ? m->mesh.nodes.data() : nullptr; }

const int64_t* tg_gmsh_node_tags(const tg_gmsh_mesh* m) {
  return m ? m->mesh.node_tags.data() : nullptr;
}

int tg_gmsh_get_block(const tg_gmsh_mesh* m, int32_t b, tg_gmsh_block* out) {
  if (!m || !out) return tg::fail(TG_ERR_INVALID_ARGUMENT, "null mesh or output");
  if (b < 0 || b >= static_cast<int32_t>(m->mesh.blocks.size()))
    return tg::fail(TG_ERR_INVALID_ARGUMENT, "block index out of range");
  const tg::gmsh::CellBlock& cb = m->mesh.blocks[b];
  out->cell = cb.cell;
  out->degree = cb.degree;
  out->num_cells = cb.tags.size();
  out->nodes_per_cell = static_cast<uint32_t>(tg::num_lagrange_nodes(cb.cell, cb.degree));
  out->nodes = cb.nodes.data();
  out->tags = cb.tags.data();
  out->elementary = cb.elementary.data();
  out->num_physical = cb.physical.size();
  out->physical = cb.physical.empty() ? nullptr : cb.physical[0].data();
  return TG_OK;
}

int tg_gmsh_get_tag_group(const tg_gmsh_mesh* m, int32_t g, tg_gmsh_tag_group* out) {
  if (!m || !out) return tg::fail(TG_ERR_INVALID_ARGUMENT, "null mesh or output");
  if (g < 0 || g >= static_cast<int32_t>(m->mesh.tag_groups.size()))
    return tg::fail(TG_ERR_INVALID_ARGUMENT, "tag group index out of range");
  const tg::gmsh::TagGroup& tg_ = m->mesh.tag_groups[g];
  out->dim = tg_.dim;
  out->cell = tg_.cell;
  out->num_elements = tg_.tags.size();
  out->corners_per_element = static_cast<uint32_t>(tg_.cell == tg::kPoint ? 1 : tg::num_vertices(tg_.cell));
  out->corners = tg_.corners.data();
  out->tags = tg_.tags.data();
  out->elementary = tg_.elementary.data();
  out->num_physical = tg_.physical.size();
  out->physical = tg_.physical.empty() ? nullptr : tg_.physical[0].data();
  return TG_OK;
}

int tg_gmsh_get_name(const tg_gmsh_mesh* m, int32_t i, int32_t* dim, int32_t* tag,
                     const char** name) {
  if (!m || !dim || !tag || !name) return tg::fail(TG_ERR_INVALID_ARGUMENT, "null argument");
  if (i < 0 || i >= static_cast<int32_t>(m->mesh.names.size()))
    return tg::fail(TG_ERR_INVALID_ARGUMENT, "name index out of range");
  *dim = m->mesh.names[i].dim;
  *tag = m->mesh.names[i].tag;
  *name = m->mesh.names[i].name.c_str();
  return TG_OK;
}

int tg_gmsh_to_internal(int32_t cell, uint32_t degree, int32_t* map, size_t cap, size_t* len) {
  if (!tg::valid_cell(cell)) return tg::fail(TG_ERR_INVALID_ARGUMENT, "invalid cell type");
  if (degree == 0) return tg::fail(TG_ERR_UNSUPPORTED_DEGREE, "UnsupportedDegree: degree 0");
  const std::vector<int>& m = tg::gmsh::gmsh_to_internal(static_cast<tg::Cell>(cell), degree);
  if (m.empty()) return tg::fail(TG_ERR_UNSUPPORTED_DEGREE, "UnsupportedDegree");
  if (len) *len = m.size();
  if (map) {
    if (cap < m.size()) return tg::fail(TG_ERR_LENGTH_MISMATCH, "map capacity too small");
    for (size_t k = 0; k < m.size(); ++k) map[k] = m[k];
  }
  return TG_OK;
}

int tg_gmsh_get_element_block(const tg_gmsh_mesh* m, int32_t i, tg_gmsh_element_block* out) {
  if (!m || !out) return tg::fail(TG_ERR_INVALID_ARGUMENT, "null mesh or output");
  if (i < 0 || i >= static_cast<int32_t>(m->element_blocks.size()))
    return tg::fail(TG_ERR_INVALID_ARGUMENT, "element block index out of range");
  *out = m->element_blocks[i];
  return TG_OK;
}

}  // extern "C"
