// VTU files on the host: the unstructured-grid arrays of a mesh (tetraloma-io/src/vtk/convert.rs:
// 116, to_vtu) and VTK's XML writer for them (write.rs:378-477, VtuFile::write_to), with the C-ABI
// entry points that take host arrays.
//
// - Points: every node, padded to three coordinates, Float64 or Float32 (rounded to nearest);
//   nodes no cell uses are written too.
// - Cells: block by block; a Lagrange cell (degree >= 2, unless linear output is asked for) lists
//   its nodes in VTK's order, a linear one its corners (pyramids of degree >= 2 are written from
//   their vertices as VTK has no Lagrange pyramid, and counted); offsets are cumulative ends.
// - The file: version 2.2, little-endian, UInt64 headers; the data either appended raw after the
//   XML or base64 inside each array, optionally zlib-compressed in blocks of 32 KiB.
#include <zlib.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "../../include/tetraloma_geom.h"
#include "vtk_internal.hpp"

namespace tg {
int fail(int code, const std::string& msg);
}

namespace tg::vtk {

namespace {

constexpr size_t kBlock = 32768;

size_t k_of(Cell c, unsigned p) { return c == kPoint ? 1 : num_lagrange_nodes(c, p); }

void put_u64(std::vector<unsigned char>& out, uint64_t v) {
  for (int i = 0; i < 8; ++i) out.push_back(static_cast<unsigned char>(v >> (8 * i)));
}

template <typename T>
Array make_array(const std::string& name, const std::string& type, uint32_t comps,
                 const std::vector<T>& v) {
  Array a;
  a.name = name;
  a.type = type;
  a.components = comps;
  a.count = v.size();
  a.bytes.resize(v.size() * sizeof(T));
  if (!v.empty()) std::memcpy(a.bytes.data(), v.data(), a.bytes.size());
  return a;
}

std::string escape(const std::string& s) {
  std::string out;
  for (char c : s) {
    switch (c) {
      case '&': out += "&amp;"; break;
      case '<': out += "&lt;"; break;
      case '>': out += "&gt;"; break;
      case '"': out += "&quot;"; break;
      case '\'': out += "&apos;"; break;
      case '\t': out += "&#9;"; break;
      case '\n': out += "&#10;"; break;
      case '\r': out += "&#13;"; break;
      default: out += c;
    }
  }
  return out;
}

void encode_base64(const unsigned char* p, size_t n, std::string& out) {
  static const char* kAlphabet =
      "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
  size_t i = 0;
  for (; i + 3 <= n; i += 3) {
    const uint32_t v = (p[i] << 16) | (p[i + 1] << 8) | p[i + 2];
    out += kAlphabet[v >> 18];
    out += kAlphabet[(v >> 12) & 63];
    out += kAlphabet[(v >> 6) & 63];
    out += kAlphabet[v & 63];
  }
  if (n - i == 1) {
    const uint32_t v = p[i] << 16;
    out += kAlphabet[v >> 18];
    out += kAlphabet[(v >> 12) & 63];
    out += "==";
  } else if (n - i == 2) {
    const uint32_t v = (p[i] << 16) | (p[i + 1] << 8);
    out += kAlphabet[v >> 18];
    out += kAlphabet[(v >> 12) & 63];
    out += kAlphabet[(v >> 6) & 63];
    out += '=';
  }
}

// The header words and the (compressed) data of one array.
void store(const std::vector<unsigned char>& raw, bool zlib, std::vector<unsigned char>& header,
           std::vector<unsigned char>& data) {
  header.clear();
  data.clear();
  if (!zlib) {
    put_u64(header, raw.size());
    header.insert(header.end(), raw.begin(), raw.end());
    return;
  }
  const size_t blocks = (raw.size() + kBlock - 1) / kBlock;
  put_u64(header, blocks);
  put_u64(header, kBlock);
  put_u64(header, raw.size() % kBlock);
  std::vector<unsigned char> buf(compressBound(kBlock));
  for (size_t b = 0; b < blocks; ++b) {
    const size_t n = std::min(kBlock, raw.size() - b * kBlock);
    uLongf len = static_cast<uLongf>(buf.size());
    if (compress2(buf.data(), &len, raw.data() + b * kBlock, static_cast<uLong>(n),
                  Z_DEFAULT_COMPRESSION) != Z_OK)
      throw Error{TG_ERR_IO, "Io: zlib compression failed"};
    data.insert(data.end(), buf.begin(), buf.begin() + len);
    put_u64(header, len);
  }
}

struct Out {
  std::string xml;
  std::vector<unsigned char> appended;
  bool base64, zlib;

  void line(int indent, const std::string& s) {
    xml.append(2 * indent, ' ');
    xml += s;
    xml += '\n';
  }

  void array(int indent, const Array& a) {
    std::vector<unsigned char> header, data;
    store(a.bytes, zlib, header, data);
    const std::string open = "<DataArray type=\"" + a.type + "\" Name=\"" + escape(a.name) +
                             "\" NumberOfComponents=\"" + std::to_string(a.components) +
                             "\" format=";
    if (!base64) {
      line(indent, open + "\"appended\" offset=\"" + std::to_string(appended.size()) + "\"/>");
      appended.insert(appended.end(), header.begin(), header.end());
      appended.insert(appended.end(), data.begin(), data.end());
    } else {
      line(indent, open + "\"binary\">");
      std::string text;
      encode_base64(header.data(), header.size(), text);
      if (zlib) encode_base64(data.data(), data.size(), text);
      line(indent + 1, text);
      line(indent, "</DataArray>");
    }
  }

  void section(const std::string& name, const std::vector<Array>& arrays) {
    line(3, "<" + name + ">");
    for (const Array& a : arrays) array(4, a);
    line(3, "</" + name + ">");
  }
};

void check_data(const std::vector<Array>& arrays, uint64_t entries) {
  for (const Array& a : arrays) {
    if (a.components == 0)
      throw Error{TG_ERR_LENGTH_MISMATCH, "LengthMismatch: a data array with no components"};
    if (a.count != entries * a.components)
      throw Error{TG_ERR_LENGTH_MISMATCH, "LengthMismatch: data array \"" + a.name + "\" holds " +
                                              std::to_string(a.count) + " values, expected " +
                                              std::to_string(entries * a.components)};
  }
}

}  // namespace

uint64_t to_vtu(int gdim, const double* nodes, uint64_t num_nodes, const std::vector<Block>& blocks,
                bool lagrange, bool f32_points, File& out) {
  if (gdim < 1 || gdim > 3)
    throw Error{TG_ERR_DIMENSION_MISMATCH,
                "DimensionMismatch: gdim " + std::to_string(gdim) + " is not 1-3"};
  // Shapes and indices first, so that nothing below reads out of range.
  uint64_t num_cells = 0, conn = 0;
  for (const Block& b : blocks) {
    if (b.degree == 0)
      throw Error{TG_ERR_UNSUPPORTED_DEGREE, std::string("UnsupportedDegree: degree 0")};
    const size_t k = k_of(b.cell, b.degree);
    if (b.len % k != 0)
      throw Error{TG_ERR_LENGTH_MISMATCH, "LengthMismatch: expected " +
                                              std::to_string(b.len - b.len % k) + ", found " +
                                              std::to_string(b.len)};
    for (uint64_t i = 0; i < b.len; ++i)
      if (b.nodes[i] >= num_nodes)
        throw Error{TG_ERR_INDEX_OUT_OF_RANGE, "IndexOutOfRange: node " +
                                                   std::to_string(b.nodes[i]) + " of " +
                                                   std::to_string(num_nodes)};
    const uint64_t cells = b.len / k;
    const bool pyramid_linear = b.cell == kPyramid && b.degree >= 2;
    const bool linear = !lagrange || b.degree == 1 || pyramid_linear;
    const size_t written = linear ? static_cast<size_t>(b.cell == kPoint ? 1 : num_vertices(b.cell)) : k;
    num_cells += cells;
    conn += cells * written;
  }
  std::vector<double> p64;
  std::vector<float> p32;
  if (f32_points)
    p32.assign(3 * num_nodes, 0.0f);
  else
    p64.assign(3 * num_nodes, 0.0);
  for (uint64_t n = 0; n < num_nodes; ++n)
    for (int g = 0; g < gdim; ++g) {
      if (f32_points)
        p32[3 * n + g] = static_cast<float>(nodes[n * gdim + g]);
      else
        p64[3 * n + g] = nodes[n * gdim + g];
    }
  out.points = f32_points ? make_array("Points", "Float32", 3, p32)
                          : make_array("Points", "Float64", 3, p64);
  out.connectivity.clear();
  out.connectivity.reserve(conn);
  out.offsets.clear();
  out.offsets.reserve(num_cells);
  out.types.clear();
  out.types.reserve(num_cells);
  uint64_t pyramids = 0;
  for (const Block& b : blocks) {
    const size_t k = k_of(b.cell, b.degree);
    const uint64_t cells = b.len / k;
    const bool pyramid_linear = b.cell == kPyramid && b.degree >= 2;
    const bool linear = !lagrange || b.degree == 1 || pyramid_linear;
    const std::vector<int>& map = vtk_to_internal(b.cell, linear ? 1 : b.degree);
    if (map.empty())
      throw Error{TG_ERR_UNSUPPORTED_DEGREE, "UnsupportedDegree: no VTK node order"};
    const uint8_t type = cell_type(b.cell, b.degree, lagrange);
    if (lagrange && pyramid_linear) pyramids += cells;
    for (uint64_t c = 0; c < cells; ++c) {
      for (int i : map) out.connectivity.push_back(b.nodes[c * k + i]);
      out.offsets.push_back(static_cast<int64_t>(out.connectivity.size()));
      out.types.push_back(type);
    }
  }
  return pyramids;
}

std::vector<unsigned char> write(const File& f, bool base64, bool zlib) {
  if (f.points.count % 3 != 0)
    throw Error{TG_ERR_LENGTH_MISMATCH, "LengthMismatch: points are not triples"};
  if (f.offsets.size() != f.types.size())
    throw Error{TG_ERR_LENGTH_MISMATCH, "LengthMismatch: offsets and types differ in length"};
  int64_t last = 0;
  for (int64_t o : f.offsets) {
    if (o < last) throw Error{TG_ERR_LENGTH_MISMATCH, "LengthMismatch: decreasing offsets"};
    last = o;
  }
  if (static_cast<uint64_t>(last) != f.connectivity.size())
    throw Error{TG_ERR_LENGTH_MISMATCH, "LengthMismatch: offsets do not end at the connectivity"};
  const uint64_t points = f.points.count / 3;
  check_data(f.point_data, points);
  check_data(f.cell_data, f.types.size());
  Out o{std::string(), {}, base64, zlib};
  o.line(0, "<?xml version=\"1.0\"?>");
  o.line(0, std::string("<VTKFile type=\"UnstructuredGrid\" version=\"2.2\" "
                        "byte_order=\"LittleEndian\" header_type=\"UInt64\"") +
                (zlib ? " compressor=\"vtkZLibDataCompressor\"" : "") + ">");
  o.line(1, "<UnstructuredGrid>");
  o.line(2, "<Piece NumberOfPoints=\"" + std::to_string(points) + "\" NumberOfCells=\"" +
                std::to_string(f.types.size()) + "\">");
  o.section("PointData", f.point_data);
  o.section("CellData", f.cell_data);
  o.line(3, "<Points>");
  o.array(4, f.points);
  o.line(3, "</Points>");
  o.line(3, "<Cells>");
  o.array(4, make_array("connectivity", "Int64", 1, f.connectivity));
  o.array(4, make_array("offsets", "Int64", 1, f.offsets));
  o.array(4, make_array("types", "UInt8", 1, f.types));
  o.line(3, "</Cells>");
  o.line(2, "</Piece>");
  o.line(1, "</UnstructuredGrid>");
  std::vector<unsigned char> bytes(o.xml.begin(), o.xml.end());
  if (!base64) {
    const std::string head = "  <AppendedData encoding=\"raw\">\n   _";
    const std::string tail = "\n  </AppendedData>\n";
    bytes.insert(bytes.end(), head.begin(), head.end());
    bytes.insert(bytes.end(), o.appended.begin(), o.appended.end());
    bytes.insert(bytes.end(), tail.begin(), tail.end());
  }
  const std::string end = "</VTKFile>\n";
  bytes.insert(bytes.end(), end.begin(), end.end());
  return bytes;
}

}  // namespace tg::vtk

namespace {

// A host data array of the C ABI as a file array.
tg::vtk::Array from_desc(const tg_vtu_array& d, uint64_t entries) {
  static const char* names[] = {"Float32", "Float64", "Int32", "UInt32", "Int64", "UInt8"};
  static const size_t sizes[] = {4, 8, 4, 4, 8, 1};
  if (d.type < 0 || d.type > 5 || !d.name)
    throw tg::vtk::Error{TG_ERR_INVALID_ARGUMENT, "bad data array type or name"};
  tg::vtk::Array a;
  a.name = d.name;
  a.type = names[d.type];
  a.components = d.components;
  a.count = d.count;
  if (d.count && !d.values) throw tg::vtk::Error{TG_ERR_INVALID_ARGUMENT, "null data values"};
  a.bytes.resize(d.count * sizes[d.type]);
  if (d.count) std::memcpy(a.bytes.data(), d.values, a.bytes.size());
  (void)entries;
  return a;
}

}  // namespace

namespace tg::vtk {

int write_file(const char* path, int gdim, const double* nodes, uint64_t num_nodes,
               const std::vector<Block>& blocks, uint32_t options, const tg_vtu_array* point_data,
               int32_t num_point_data, const tg_vtu_array* cell_data, int32_t num_cell_data,
               uint64_t* pyramids) {
  try {
    File f;
    const uint64_t pyr = to_vtu(gdim, nodes, num_nodes, blocks, !(options & TG_VTU_LINEAR),
                                (options & TG_VTU_F32_POINTS) != 0, f);
    for (int32_t i = 0; i < num_point_data; ++i) f.point_data.push_back(from_desc(point_data[i], 0));
    for (int32_t i = 0; i < num_cell_data; ++i) f.cell_data.push_back(from_desc(cell_data[i], 0));
    const std::vector<unsigned char> bytes =
        write(f, (options & TG_VTU_BASE64) != 0, (options & TG_VTU_ZLIB) != 0);
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return tg::fail(TG_ERR_IO, std::string("Io: ") + path + ": cannot create");
    const bool ok = std::fwrite(bytes.data(), 1, bytes.size(), fp) == bytes.size();
    const bool closed = std::fclose(fp) == 0;
    if (!ok || !closed) return tg::fail(TG_ERR_IO, std::string("Io: ") + path + ": write error");
    if (pyramids) *pyramids = pyr;
    return TG_OK;
  } catch (const Error& e) {
    return tg::fail(e.code, e.message);
  } catch (const std::bad_alloc&) {
    return tg::fail(TG_ERR_TOO_LARGE, "TooLarge: out of host memory writing the VTU file");
  }
}

}  // namespace tg::vtk

extern "C" {

int tg_vtk_to_internal(int32_t cell, uint32_t degree, int32_t* map, size_t cap, size_t* len) {
  if (!tg::valid_cell(cell)) return tg::fail(TG_ERR_INVALID_ARGUMENT, "invalid cell type");
  if (degree == 0) return tg::fail(TG_ERR_UNSUPPORTED_DEGREE, "UnsupportedDegree: degree 0");
  const std::vector<int>& m = tg::vtk::vtk_to_internal(static_cast<tg::Cell>(cell), degree);
  if (m.empty())
    return tg::fail(TG_ERR_UNSUPPORTED_DEGREE,
                    "UnsupportedDegree: VTK has no Lagrange cell of this type and degree");
  if (len) *len = m.size();
  if (map) {
    if (cap < m.size()) return tg::fail(TG_ERR_LENGTH_MISMATCH, "map capacity too small");
    for (size_t k = 0; k < m.size(); ++k) map[k] = m[k];
  }
  return TG_OK;
}

int32_t tg_vtk_cell_type(int32_t cell, uint32_t degree, int32_t lagrange) {
  if (!tg::valid_cell(cell) || degree == 0) return -1;
  return tg::vtk::cell_type(static_cast<tg::Cell>(cell), degree, lagrange != 0);
}

int tg_vtu_write_mesh(const char* path, int32_t gdim, const double* nodes, uint64_t num_nodes,
                      const tg_vtu_block* blocks, int32_t num_blocks, uint32_t options,
                      const tg_vtu_array* point_data, int32_t num_point_data,
                      const tg_vtu_array* cell_data, int32_t num_cell_data, uint64_t* pyramids) {
  if (!path || (num_nodes && !nodes) || (num_blocks && !blocks) || num_blocks < 0 ||
      num_point_data < 0 || num_cell_data < 0 || (num_point_data && !point_data) ||
      (num_cell_data && !cell_data))
    return tg::fail(TG_ERR_INVALID_ARGUMENT, "null or negative argument");
  std::vector<tg::vtk::Block> bl;
  for (int32_t b = 0; b < num_blocks; ++b) {
    if (!tg::valid_cell(blocks[b].cell) || (blocks[b].len && !blocks[b].nodes))
      return tg::fail(TG_ERR_INVALID_ARGUMENT, "bad block");
    bl.push_back({static_cast<tg::Cell>(blocks[b].cell), blocks[b].degree, blocks[b].nodes,
                  blocks[b].len});
  }
  return tg::vtk::write_file(path, gdim, nodes, num_nodes, bl, options, point_data,
                             num_point_data, cell_data, num_cell_data, pyramids);
}

}  // extern "C"
