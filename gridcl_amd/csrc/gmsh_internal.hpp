// Gmsh MSH 4.1 reading on the host: the file's sections as stored, the element table, Gmsh's
// node order for Lagrange elements, and the mesh a file describes (cell blocks in A.5 node order,
// lower-dimensional elements as tag groups). Follows tetraloma-io/src/gmsh (model.rs, parse.rs,
// element_type.rs, chunk.rs, convert.rs) and tetraloma-refcell/src/node_maps/gmsh.rs,
// serendipity.rs.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "refcell.hpp"

namespace tg {
int fail(int code, const std::string& msg);
const char* cell_name(int c);
}

namespace tg::gmsh {

// An error of the reader: a tg_status code, its message and its detail (the byte offset of a
// Parse error, the repeated tag of DuplicateOriginId, the element of UnmatchedTagElement).
struct Error {
  int code;
  std::string message;
  int64_t detail = 0;
};

// ---- The file model (model.rs) ------------------------------------------------------------

struct PhysicalName {
  int dim;
  int32_t tag;
  std::string name;
};

struct Entity {
  int32_t tag = 0;
  double min[3] = {0, 0, 0}, max[3] = {0, 0, 0};
  std::vector<int32_t> physical, boundary;
};

struct PartitionedEntity {
  Entity entity;
  int parent_dim = 0;
  int32_t parent_tag = 0;
  std::vector<int32_t> partitions;
};

struct Periodic {
  int dim;
  int32_t tag, master_tag;
  std::vector<double> affine;
  std::vector<std::array<int64_t, 2>> nodes;
};

struct NodeBlock {
  int dim;
  int32_t entity_tag;
  std::vector<int64_t> tags;
  std::vector<double> coords;  // [node][3]
};

struct ElementBlock {
  int dim;
  int32_t entity_tag;
  int32_t element_type;
  std::vector<int64_t> tags;
  std::vector<int64_t> nodes;  // [element][nodes per element], Gmsh order
};

struct File {
  bool binary = false;
  std::vector<PhysicalName> names;
  bool has_entities = false;
  std::vector<Entity> entities[4];
  bool has_partitioned = false;
  uint64_t num_partitions = 0;
  std::vector<std::array<int32_t, 2>> ghost_entities;
  std::vector<PartitionedEntity> partitioned[4];
  std::vector<Periodic> periodic;
  std::vector<NodeBlock> node_blocks;
  std::vector<ElementBlock> element_blocks;
};

// Parses an MSH 4.1 file (ASCII or binary, either byte order, 4- or 8-byte size_t); throws Error.
File parse(const unsigned char* bytes, size_t len);

// ---- The element table (element_type.rs) --------------------------------------------------

struct ElementType {
  int32_t id;
  Cell cell;
  unsigned degree;  // 2 for the serendipity types, 1 for the point
  size_t num_nodes;
  bool serendipity;
};

// The type `id`, or throws UnsupportedElementType.
ElementType element_type(int32_t id);
// The complete Lagrange type of (cell, degree), or throws UnsupportedDegree.
ElementType lagrange_element_type(Cell cell, unsigned degree);

// ---- Node orders (node_maps/gmsh.rs, serendipity.rs) ---------------------------------------

// map[k] = the A.5 index of Gmsh node k of the degree-p Lagrange element on `cell` (p >= 1).
const std::vector<int>& gmsh_to_internal(Cell cell, unsigned p);
// The weights [missing][file node] that give the missing complete-P2 nodes (Gmsh P2 indices
// n..m-1) from the n nodes of Gmsh's serendipity element on `cell`.
const std::vector<double>& serendipity_weights(Cell cell, size_t* num_file_nodes,
                                               size_t* num_nodes);

// ---- The mesh (convert.rs) -----------------------------------------------------------------

struct CellBlock {
  Cell cell;
  unsigned degree;
  std::vector<uint32_t> nodes;  // [cell][k], A.5 order, indices into Mesh::nodes
  std::vector<int64_t> tags;
  std::vector<int32_t> elementary;
  std::vector<std::array<int64_t, 2>> physical;  // (cell index, physical tag), sorted
};

struct TagGroup {
  int dim;
  Cell cell;
  std::vector<int64_t> corners;  // [element][vertices], internal vertex order
  std::vector<int64_t> tags;
  std::vector<int32_t> elementary;
  std::vector<std::array<int64_t, 2>> physical;
};

struct Mesh {
  int tdim = 0, gdim = 0;
  std::vector<double> nodes;  // [node][gdim]
  std::vector<int64_t> node_tags;
  std::vector<CellBlock> blocks;
  std::vector<TagGroup> tag_groups;
  std::vector<PhysicalName> names;
};

// The mesh `file` describes; gdim 0 infers it from the coordinates. Throws Error.
Mesh to_mesh(const File& file, int gdim);

}  // namespace tg::gmsh
