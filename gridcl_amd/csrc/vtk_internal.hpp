// VTU output on the host: VTK's Lagrange node order, the unstructured-grid arrays of a mesh and
// the XML file writer. Follows tetraloma-io/src/vtk (cell_type.rs, convert.rs, write.rs,
// base64.rs) and tetraloma-refcell/src/node_maps/vtk.rs.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "refcell.hpp"

namespace tg::vtk {

struct Error {
  int code;
  std::string message;
};

// map[k] = the A.5 index of VTK's k-th node of the degree-p Lagrange cell (p = 1 for the
// pyramid); empty if the format has no such cell.
const std::vector<int>& vtk_to_internal(Cell cell, unsigned p);

// The VTK cell type written for (cell, degree) (cell_type.rs:41-60).
uint8_t cell_type(Cell cell, unsigned degree, bool lagrange);

struct Block {
  Cell cell;
  unsigned degree;
  const uint32_t* nodes;  // [cells][k], A.5 order
  uint64_t len;           // entries in nodes
};

// A data array of the file: VTK type name, components and little-endian values.
struct Array {
  std::string name;
  std::string type;  // Float32, Float64, Int32, UInt32, Int64, UInt8
  uint32_t components = 1;
  std::vector<unsigned char> bytes;
  uint64_t count = 0;  // values
};

struct File {
  Array points;  // Float32 or Float64, 3 components
  std::vector<int64_t> connectivity, offsets;
  std::vector<uint8_t> types;
  std::vector<Array> point_data, cell_data;
};

// The file of a mesh ([node][gdim] f64 coordinates, blocks): convert.rs:116 (to_vtu). Returns
// the number of pyramids of degree >= 2 written from their vertices. Throws Error.
uint64_t to_vtu(int gdim, const double* nodes, uint64_t num_nodes, const std::vector<Block>& blocks,
                bool lagrange, bool f32_points, File& out);

// The XML bytes of `file`: raw data appended after the XML or base64 inside each array,
// optionally zlib-compressed in 32 KiB blocks (write.rs:378-477). Throws Error.
std::vector<unsigned char> write(const File& file, bool base64, bool zlib);

}  // namespace tg::vtk
