// Reading the sections of an MSH 4.1 file (the format tetraloma-io/src/gmsh/parse.rs reads) into
// the file model of gmsh_internal.hpp: text or binary data, either byte order, 4- or 8-byte
// size_t.
//
// Design. The format writes every section with three kinds of numbers (int, size_t, double),
// as text tokens or as fixed-width binary values. Two number decoders with the same interface
// (TextNumbers, BinaryNumbers) are chosen once from $MeshFormat, and every section decoder is
// written once as a template over the decoder, so the text and binary paths cannot drift apart.
// Binary arrays of tags and coordinates are converted in bulk. Records shared between sections
// are decoded by shared helpers: an entity's bounding box and tag lists ($Entities and
// $PartitionedEntities), the four-number block header and the announced totals ($Nodes and
// $Elements). The sections are dispatched through a table; names not in it are skipped to their
// $End line.
//
// Safety. Every count is checked against the bytes left (each value takes at least one byte and a
// separator in text, its width in binary) before anything is sized from it, so a damaged file is a
// Parse error with the byte offset where it was found, never a crash or a large allocation.
#include <charconv>
#include <cstring>
#include <limits>
#include <string>
#include <string_view>

#include "gmsh_internal.hpp"

namespace tg::gmsh {

namespace {

constexpr int kParse = 15;  // TG_ERR_PARSE

[[noreturn]] void parse_error(size_t offset, const std::string& what) {
  throw Error{kParse, "Parse: byte " + std::to_string(offset) + " of the Gmsh file: " + what,
              static_cast<int64_t>(offset)};
}

bool blank(unsigned char c) { return c == ' ' || c == '\n' || c == '\r' || c == '\t' || c == '\f'; }

// A file's word quoted in a message: printable ASCII as is, other bytes as \xNN, at most 40
// bytes of it (a damaged file may hold anything).
std::string quoted(std::string_view w) {
  static const char hex[] = "0123456789abcdef";
  std::string out = "\"";
  for (size_t i = 0; i < w.size() && i < 40; ++i) {
    const unsigned char c = static_cast<unsigned char>(w[i]);
    if (c >= 0x20 && c < 0x7f && c != '\\' && c != '"') {
      out.push_back(static_cast<char>(c));
    } else {
      out += "\\x";
      out.push_back(hex[c >> 4]);
      out.push_back(hex[c & 15]);
    }
  }
  if (w.size() > 40) out += "...";
  return out + "\"";
}

// The file's bytes and the read position.
struct Bytes {
  const unsigned char* data;
  size_t size;
  size_t at = 0;

  size_t left() const { return size - at; }
  void skip_blanks() {
    while (at < size && blank(data[at])) ++at;
  }
  // The next blank-separated word, and where it starts.
  std::string_view word(size_t* start) {
    skip_blanks();
    *start = at;
    while (at < size && !blank(data[at])) ++at;
    if (*start == at) parse_error(at, "the file ends before the next value");
    return {reinterpret_cast<const char*>(data + *start), at - *start};
  }
  const unsigned char* consume(size_t n) {
    if (left() < n) parse_error(at, "the file ends " + std::to_string(n - left()) +
                                        " bytes early");
    const unsigned char* p = data + at;
    at += n;
    return p;
  }
};

// Tags (node, element, periodic node) are size_t values from 1 to 2^63 - 1.
int64_t checked_tag(uint64_t v, size_t offset, const char* kind) {
  if (v == 0 || v > static_cast<uint64_t>(std::numeric_limits<int64_t>::max()))
    parse_error(offset, std::string(kind) + " tag " + std::to_string(v) + " is outside 1..2^63-1");
  return static_cast<int64_t>(v);
}

// Numbers written as text tokens.
class TextNumbers {
 public:
  explicit TextNumbers(Bytes& b) : b_(b) {}
  static constexpr bool kBinary = false;
  // The fewest bytes an int, size_t or double takes: one character and a separator.
  size_t int_bytes() const { return 2; }
  size_t size_bytes() const { return 2; }
  size_t real_bytes() const { return 2; }
  Bytes& bytes() { return b_; }

  int32_t i32() { return number<int32_t>("an int"); }
  uint64_t word() { return number<uint64_t>("a size_t"); }
  double real() { return number<double>("a double"); }
  void reals(double* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = real();
  }
  // n tags, written to out[i * stride + offset_of(i)] by `place`.
  template <class Place>
  void tags(size_t n, const char* kind, Place place) {
    for (size_t i = 0; i < n; ++i) {
      const size_t at = next_offset();
      place(i, checked_tag(word(), at, kind));
    }
  }
  size_t next_offset() {
    b_.skip_blanks();
    return b_.at;
  }
  void section_body() {}  // text data may follow the header on the same line

 private:
  template <typename T>
  T number(const char* what) {
    size_t start;
    std::string_view w = b_.word(&start);
    std::string_view digits = w;
    if (digits.size() > 1 && digits[0] == '+' && digits[1] != '-' && digits[1] != '+')
      digits.remove_prefix(1);
    T v{};
    const auto res = std::from_chars(digits.data(), digits.data() + digits.size(), v);
    if (res.ec != std::errc() || res.ptr != digits.data() + digits.size())
      parse_error(start, quoted(w) + " is not " + what);
    return v;
  }
  Bytes& b_;
};

// Numbers written as binary values of fixed width (int 4 bytes, size_t 4 or 8, double 8).
class BinaryNumbers {
 public:
  BinaryNumbers(Bytes& b, bool big_endian, size_t size_width)
      : b_(b), big_(big_endian), sw_(size_width) {}
  static constexpr bool kBinary = true;
  size_t int_bytes() const { return 4; }
  size_t size_bytes() const { return sw_; }
  size_t real_bytes() const { return 8; }
  Bytes& bytes() { return b_; }

  int32_t i32() { return static_cast<int32_t>(static_cast<uint32_t>(load(b_.consume(4), 4))); }
  uint64_t word() { return load(b_.consume(sw_), sw_); }
  double real() {
    const uint64_t bits = load(b_.consume(8), 8);
    double d;
    std::memcpy(&d, &bits, 8);
    return d;
  }
  void reals(double* out, size_t n) {
    if (n == 0) return;  // (out may be the null data of an empty vector)
    const unsigned char* p = b_.consume(n * 8);
    if (!big_) {
      std::memcpy(out, p, n * 8);
      return;
    }
    for (size_t i = 0; i < n; ++i) {
      const uint64_t bits = load(p + 8 * i, 8);
      std::memcpy(out + i, &bits, 8);
    }
  }
  template <class Place>
  void tags(size_t n, const char* kind, Place place) {
    const size_t base = b_.at;
    const unsigned char* p = b_.consume(n * sw_);
    for (size_t i = 0; i < n; ++i) place(i, checked_tag(load(p + i * sw_, sw_), base + i * sw_, kind));
  }
  size_t next_offset() { return b_.at; }
  // Binary data starts right after the line break that ends the section's header line.
  void section_body() {
    if (b_.at < b_.size && b_.data[b_.at] == '\r') ++b_.at;
    if (b_.at >= b_.size || b_.data[b_.at] != '\n')
      parse_error(b_.at, "binary data must start on the line after the section name");
    ++b_.at;
  }

 private:
  uint64_t load(const unsigned char* p, size_t w) const {
    uint64_t v = 0;
    if (big_) {
      for (size_t i = 0; i < w; ++i) v = (v << 8) | p[i];
    } else if (w == 8) {
      std::memcpy(&v, p, 8);
    } else {
      uint32_t u;
      std::memcpy(&u, p, 4);
      v = u;
    }
    return v;
  }
  Bytes& b_;
  bool big_;
  size_t sw_;
};

// A count that the rest of the file can hold at `per_item` bytes per item at least.
template <class Num>
size_t bounded_count(Num& in, size_t per_item) {
  const size_t at = in.next_offset();
  const uint64_t n = in.word();
  const size_t left = in.bytes().left();
  if (per_item > 0 && n > left / per_item)
    parse_error(at, "a count of " + std::to_string(n) + " does not fit in the " +
                        std::to_string(left) + " bytes left");
  return static_cast<size_t>(n);
}

template <class Num>
int dimension(Num& in) {
  const size_t at = in.next_offset();
  const int32_t d = in.i32();
  if (d < 0 || d > 3) parse_error(at, "dimension " + std::to_string(d) + " (0 to 3 expected)");
  return d;
}

// A counted list of ints (physical tags, bounding entities, partitions).
template <class Num>
std::vector<int32_t> int_list(Num& in) {
  std::vector<int32_t> out(bounded_count(in, in.int_bytes()));
  for (int32_t& v : out) v = in.i32();
  return out;
}

// The part of an entity record after its tag (and, partitioned, its parent and partitions):
// the point (dimension 0) or bounding box, the physical tags and, above dimension 0, the
// bounding entities.
template <class Num>
void entity_rest(Num& in, int dim, Entity& e) {
  in.reals(e.min, 3);
  if (dim == 0)
    std::memcpy(e.max, e.min, sizeof e.min);
  else
    in.reals(e.max, 3);
  e.physical = int_list(in);
  if (dim > 0) e.boundary = int_list(in);
}

// The four per-dimension entity counts that open $Entities and $PartitionedEntities.
template <class Num>
void dimension_counts(Num& in, size_t min_record, size_t (&counts)[4]) {
  for (size_t& c : counts) c = bounded_count(in, min_record);
}

template <class Num>
void read_entities(Num& in, File& f) {
  size_t counts[4];
  dimension_counts(in, in.int_bytes() + 3 * in.real_bytes(), counts);
  for (int dim = 0; dim < 4; ++dim) {
    std::vector<Entity>& list = f.entities[dim];
    list.assign(counts[dim], Entity{});
    for (Entity& e : list) {
      e.tag = in.i32();
      entity_rest(in, dim, e);
    }
  }
  f.has_entities = true;
}

template <class Num>
void read_partitioned_entities(Num& in, File& f) {
  f.num_partitions = in.word();
  f.ghost_entities.assign(bounded_count(in, 2 * in.int_bytes()), {0, 0});
  for (auto& g : f.ghost_entities) {
    g[0] = in.i32();
    g[1] = in.i32();
  }
  size_t counts[4];
  dimension_counts(in, 3 * in.int_bytes() + 3 * in.real_bytes(), counts);
  for (int dim = 0; dim < 4; ++dim) {
    std::vector<PartitionedEntity>& list = f.partitioned[dim];
    list.assign(counts[dim], PartitionedEntity{});
    for (PartitionedEntity& pe : list) {
      pe.entity.tag = in.i32();
      pe.parent_dim = dimension(in);
      pe.parent_tag = in.i32();
      pe.partitions = int_list(in);
      entity_rest(in, dim, pe.entity);
    }
  }
  f.has_partitioned = true;
}

template <class Num>
void read_periodic(Num& in, File& f) {
  const size_t links = bounded_count(in, 3 * in.int_bytes() + 2 * in.size_bytes());
  for (size_t l = 0; l < links; ++l) {
    Periodic per;
    per.dim = dimension(in);
    per.tag = in.i32();
    per.master_tag = in.i32();
    per.affine.resize(bounded_count(in, in.real_bytes()));
    in.reals(per.affine.data(), per.affine.size());
    per.nodes.resize(bounded_count(in, 2 * in.size_bytes()));
    in.tags(2 * per.nodes.size(), "periodic node",
            [&](size_t i, int64_t t) { per.nodes[i / 2][i % 2] = t; });
    f.periodic.push_back(std::move(per));
  }
}

// The header of $Nodes and $Elements (blocks, total, smallest and largest tag), and the check
// that the blocks hold the announced total.
struct Totals {
  size_t blocks = 0;
  uint64_t announced = 0, found = 0;
  size_t at = 0;
};

template <class Num>
Totals section_totals(Num& in) {
  Totals t;
  t.at = in.next_offset();
  t.blocks = bounded_count(in, 3 * in.int_bytes() + in.size_bytes());
  t.announced = in.word();
  (void)in.word();  // smallest tag
  (void)in.word();  // largest tag
  return t;
}

void check_totals(const Totals& t, const char* what) {
  if (t.found != t.announced)
    parse_error(t.at, "the header announces " + std::to_string(t.announced) + " " + what +
                          " but the blocks hold " + std::to_string(t.found));
}

template <class Num>
void read_nodes(Num& in, File& f) {
  Totals tot = section_totals(in);
  for (size_t blk = 0; blk < tot.blocks; ++blk) {
    NodeBlock nb;
    nb.dim = dimension(in);
    nb.entity_tag = in.i32();
    const size_t flag_at = in.next_offset();
    const int32_t parametric = in.i32();
    if (parametric != 0 && parametric != 1)
      parse_error(flag_at, "parametric flag " + std::to_string(parametric) + " (0 or 1 expected)");
    // Each node: x, y, z and, parametric, its dim parametric coordinates (dropped here).
    const size_t per_node = 3 + (parametric ? static_cast<size_t>(nb.dim) : 0);
    const size_t n = bounded_count(in, in.size_bytes() + per_node * in.real_bytes());
    nb.tags.resize(n);
    in.tags(n, "node", [&](size_t i, int64_t t) { nb.tags[i] = t; });
    nb.coords.resize(3 * n);
    if (per_node == 3) {
      in.reals(nb.coords.data(), 3 * n);
    } else {
      double row[6];
      for (size_t i = 0; i < n; ++i) {
        in.reals(row, per_node);
        std::memcpy(&nb.coords[3 * i], row, 3 * sizeof(double));
      }
    }
    tot.found += n;
    f.node_blocks.push_back(std::move(nb));
  }
  check_totals(tot, "nodes");
}

template <class Num>
void read_elements(Num& in, File& f) {
  Totals tot = section_totals(in);
  for (size_t blk = 0; blk < tot.blocks; ++blk) {
    ElementBlock eb;
    eb.dim = dimension(in);
    eb.entity_tag = in.i32();
    const size_t type_at = in.next_offset();
    eb.element_type = in.i32();
    const ElementType type = element_type(eb.element_type);  // UnsupportedElementType
    if (tdim(type.cell) != eb.dim)
      parse_error(type_at, "element type " + std::to_string(eb.element_type) + " has dimension " +
                               std::to_string(tdim(type.cell)) + ", its block " +
                               std::to_string(eb.dim));
    const size_t k = type.num_nodes;
    const size_t n = bounded_count(in, (k + 1) * in.size_bytes());
    eb.tags.resize(n);
    eb.nodes.resize(n * k);
    // Each row: the element's tag, then its k node tags.
    in.tags(n * (k + 1), "element or node", [&](size_t i, int64_t t) {
      const size_t row = i / (k + 1), col = i % (k + 1);
      if (col == 0)
        eb.tags[row] = t;
      else
        eb.nodes[row * k + col - 1] = t;
    });
    tot.found += n;
    f.element_blocks.push_back(std::move(eb));
  }
  check_totals(tot, "elements");
}

// $PhysicalNames is text in binary files too: dim, tag and a quoted name per line.
void read_physical_names(Bytes& b, File& f) {
  TextNumbers text(b);
  const size_t n = bounded_count(text, 3 * text.int_bytes() + 2);
  for (size_t i = 0; i < n; ++i) {
    PhysicalName pn;
    pn.dim = dimension(text);
    pn.tag = text.i32();
    b.skip_blanks();
    const size_t open = b.at;
    if (b.left() == 0 || b.data[b.at] != '"') parse_error(open, "a physical name must be quoted");
    const unsigned char* first = b.data + open + 1;
    const void* close = std::memchr(first, '"', b.size - open - 1);
    if (!close) parse_error(open, "the quoted physical name is not closed");
    const size_t len = static_cast<const unsigned char*>(close) - first;
    pn.name.assign(reinterpret_cast<const char*>(first), len);
    if (pn.name.find_first_of("\r\n") != std::string::npos)
      parse_error(open, "a physical name spans lines");
    b.at = open + len + 2;
    f.names.push_back(std::move(pn));
  }
}

// The closing line of a section.
void expect_end(Bytes& b, std::string_view name) {
  size_t start;
  const std::string_view w = b.word(&start);
  if (w.size() != name.size() + 4 || w.compare(0, 4, "$End") != 0 || w.substr(4) != name)
    parse_error(start, quoted(w) + " where " + quoted("$End" + std::string(name)) +
                           " closes the section");
}

// The data format of $MeshFormat: version, file type and size_t width, and for binary files the
// int 1 that shows the byte order.
struct Format {
  bool binary = false, big_endian = false;
  size_t size_width = 8;
};

Format read_format(Bytes& b) {
  size_t start;
  if (b.word(&start) != "$MeshFormat") parse_error(start, "a Gmsh file starts with $MeshFormat");
  const std::string_view version = b.word(&start);
  if (version != "4.1")
    parse_error(start, "version " + quoted(version) + " (this reader reads MSH 4.1)");
  TextNumbers text(b);
  Format fmt;
  const size_t type_at = text.next_offset();
  const int32_t type = text.i32();
  const int32_t width = text.i32();
  if (type != 0 && type != 1)
    parse_error(type_at, "file type " + std::to_string(type) + " (0 text, 1 binary)");
  if (width != 4 && width != 8)
    parse_error(type_at, "size_t width " + std::to_string(width) + " (4 or 8)");
  fmt.size_width = static_cast<size_t>(width);
  if (type == 1) {
    fmt.binary = true;
    BinaryNumbers(b, false, fmt.size_width).section_body();
    const size_t one_at = b.at;
    const unsigned char* one = b.consume(4);
    const unsigned char little[4] = {1, 0, 0, 0}, big[4] = {0, 0, 0, 1};
    if (std::memcmp(one, big, 4) == 0)
      fmt.big_endian = true;
    else if (std::memcmp(one, little, 4) != 0)
      parse_error(one_at, "the binary int 1 reads as neither byte order");
  }
  expect_end(b, "MeshFormat");
  return fmt;
}

// The numeric sections by name, for a decoder type.
template <class Num>
struct SectionTable {
  using Decode = void (*)(Num&, File&);
  struct Entry {
    std::string_view name;
    Decode decode;
  };
  static constexpr Entry kEntries[] = {
      {"Entities", read_entities<Num>},
      {"PartitionedEntities", read_partitioned_entities<Num>},
      {"Periodic", read_periodic<Num>},
      {"Nodes", read_nodes<Num>},
      {"Elements", read_elements<Num>},
  };
  static Decode find(std::string_view name) {
    for (const Entry& e : kEntries)
      if (e.name == name) return e.decode;
    return nullptr;
  }
};

// Every section after $MeshFormat, with the file's number decoder.
template <class Num>
void read_sections(Bytes& b, Num& in, File& f) {
  for (b.skip_blanks(); b.left() > 0; b.skip_blanks()) {
    size_t start;
    const std::string_view head = b.word(&start);
    if (head[0] != '$') parse_error(start, quoted(head) + " where a section starts");
    const std::string_view name = head.substr(1);
    if (name == "MeshFormat") parse_error(start, "$MeshFormat appears twice");
    if (name.compare(0, 3, "End") == 0)
      parse_error(start, quoted(head) + " closes a section that was not opened");
    if (name == "PhysicalNames") {
      read_physical_names(b, f);
    } else if (auto decode = SectionTable<Num>::find(name)) {
      in.section_body();
      decode(in, f);
    } else {
      // Not a section of the mesh: skip to its $End line.
      const std::string close = "$End" + std::string(name);
      const void* hit = b.left() >= close.size()
                            ? memmem(b.data + b.at, b.left(), close.data(), close.size())
                            : nullptr;
      if (!hit) parse_error(start, "section " + quoted(head) + " has no " + quoted(close));
      b.at = static_cast<const unsigned char*>(hit) - b.data;
    }
    expect_end(b, name);
  }
}

}  // namespace

File parse(const unsigned char* bytes, size_t len) {
  Bytes b{bytes, len};
  const Format fmt = read_format(b);
  File f;
  f.binary = fmt.binary;
  if (fmt.binary) {
    BinaryNumbers in(b, fmt.big_endian, fmt.size_width);
    read_sections(b, in, f);
  } else {
    TextNumbers in(b);
    read_sections(b, in, f);
  }
  return f;
}

}  // namespace tg::gmsh
