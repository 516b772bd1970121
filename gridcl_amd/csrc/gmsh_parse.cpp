// The MSH 4.1 parser (tetraloma-io/src/gmsh/parse.rs:14-620): ASCII and binary files, either
// byte order and either size_t width, into the file's sections as stored.
//
// Counts are checked against the bytes left before anything is sized from them, so a damaged file
// gives a Parse error (with the byte offset where it was found), never a crash or a huge
// allocation. Unknown sections are skipped; $PhysicalNames is text in binary files too.
#include <charconv>
#include <cstring>
#include <limits>
#include <string>
#include <string_view>

#include "gmsh_internal.hpp"

namespace tg::gmsh {

namespace {

constexpr int kParse = 15;
constexpr int kUnsupportedElementType = 16;
// The fewest bytes a text value takes: one character and one separator (parse.rs:23).
constexpr size_t kToken = 2;

bool is_space(unsigned char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f'; }

class Cursor {
 public:
  Cursor(const unsigned char* b, size_t n) : b_(b), n_(n) {}

  [[noreturn]] void error_at(size_t at, const std::string& m) const {
    throw Error{kParse, "Parse: Gmsh, byte " + std::to_string(at) + ": " + m,
                static_cast<int64_t>(at)};
  }
  [[noreturn]] void error(const std::string& m) const { error_at(pos_, m); }

  size_t pos() const { return pos_; }
  size_t remaining() const { return n_ - pos_; }
  bool binary() const { return binary_; }

  void skip_whitespace() {
    while (pos_ < n_ && is_space(b_[pos_])) ++pos_;
  }
  bool at_end() {
    skip_whitespace();
    return pos_ == n_;
  }
  std::string_view token(size_t* at) {
    skip_whitespace();
    const size_t start = pos_;
    while (pos_ < n_ && !is_space(b_[pos_])) ++pos_;
    if (start == pos_) error("unexpected end of file");
    *at = start;
    return std::string_view(reinterpret_cast<const char*>(b_ + start), pos_ - start);
  }

  template <typename T>
  T parse_text(const char* what) {
    size_t at;
    std::string_view t = token(&at);
    std::string_view digits = t;
    if (!digits.empty() && digits[0] == '+') digits.remove_prefix(1);
    T v{};
    auto r = std::from_chars(digits.data(), digits.data() + digits.size(), v);
    if (r.ec != std::errc() || r.ptr != digits.data() + digits.size() || digits.empty() ||
        (t.size() > 1 && t[0] == '+' && t[1] == '-'))
      error_at(at, std::string("expected ") + what + ", found \"" + std::string(t) + "\"");
    return v;
  }

  const unsigned char* take(size_t n) {
    if (remaining() < n) error("unexpected end of file");
    const unsigned char* p = b_ + pos_;
    pos_ += n;
    return p;
  }

  uint64_t raw(size_t width) {
    const unsigned char* p = take(width);
    uint64_t v = 0;
    if (big_)
      for (size_t i = 0; i < width; ++i) v = (v << 8) | p[i];
    else
      for (size_t i = width; i-- > 0;) v = (v << 8) | p[i];
    return v;
  }

  int32_t int_() {
    if (binary_) return static_cast<int32_t>(static_cast<uint32_t>(raw(4)));
    return parse_text<int32_t>("an integer");
  }
  uint64_t size() {
    if (binary_) return raw(size_t_);
    return parse_text<uint64_t>("an unsigned integer");
  }
  double float_() {
    if (binary_) {
      const uint64_t bits = raw(8);
      double d;
      std::memcpy(&d, &bits, 8);
      return d;
    }
    return parse_text<double>("a number");
  }

  // n tags (size_t values >= 1) into out, the i-th at out[i + i / run * skip]: binary arrays are
  // converted in one pass (run = 0: contiguous).
  void tags_into(int64_t* out, size_t n, const char* what) { tags_split(out, nullptr, n, 0, what); }

  // Rows of (1 + k) tags: the first of each row into `first`, the other k into `rest`.
  void tags_split(int64_t* rest, int64_t* first, size_t n, size_t k, const char* what) {
    const size_t per = first ? k + 1 : 1;
    auto put = [&](size_t i, int64_t v) {
      if (!first) {
        rest[i] = v;
      } else {
        const size_t r = i / per, j = i % per;
        if (j == 0)
          first[r] = v;
        else
          rest[r * k + j - 1] = v;
      }
    };
    if (!binary_) {
      for (size_t i = 0; i < n * per; ++i) put(i, tag(what));
      return;
    }
    const size_t start = pos_;
    const unsigned char* p = take(n * per * size_t_);
    for (size_t i = 0; i < n * per; ++i) {
      const unsigned char* q = p + i * size_t_;
      uint64_t v = 0;
      if (big_)
        for (size_t j = 0; j < size_t_; ++j) v = (v << 8) | q[j];
      else if (size_t_ == 8)
        std::memcpy(&v, q, 8);
      else {
        uint32_t w;
        std::memcpy(&w, q, 4);
        v = w;
      }
      if (v < 1 || v > static_cast<uint64_t>(std::numeric_limits<int64_t>::max()))
        error_at(start + i * size_t_, std::string(what) + " tag " + std::to_string(v) +
                                          ": tags must be 1 to 2^63 - 1");
      put(i, static_cast<int64_t>(v));
    }
  }

  // n doubles into out.
  void floats_into(double* out, size_t n) {
    if (!binary_) {
      for (size_t i = 0; i < n; ++i) out[i] = float_();
      return;
    }
    const unsigned char* p = take(n * 8);
    if (!big_) {
      std::memcpy(out, p, n * 8);
      return;
    }
    for (size_t i = 0; i < n; ++i) {
      uint64_t v = 0;
      for (size_t j = 0; j < 8; ++j) v = (v << 8) | p[i * 8 + j];
      std::memcpy(&out[i], &v, 8);
    }
  }

  // The line break after a section header; binary data starts after it.
  void header_end() {
    if (!binary_) return;
    if (pos_ < n_ && b_[pos_] == '\r') ++pos_;
    if (pos_ >= n_ || b_[pos_] != '\n') error("expected a line break after the section header");
    ++pos_;
  }

  int dim() {
    const size_t at = pos_;
    const int32_t d = int_();
    if (d < 0 || d > 3) error_at(at, "dimension " + std::to_string(d) + " is not 0-3");
    return d;
  }

  // A count of items of at least `min_bytes` bytes each that the rest of the file can hold.
  size_t count(size_t min_bytes) {
    const size_t at = pos_;
    const uint64_t n = size();
    const bool fits = min_bytes == 0 ? true
                                     : n <= remaining() / min_bytes;
    if (!fits) error_at(at, "count " + std::to_string(n) + " exceeds the rest of the file");
    return static_cast<size_t>(n);
  }

  int64_t tag(const char* what) {
    const size_t at = pos_;
    const uint64_t t = size();
    if (t < 1 || t > static_cast<uint64_t>(std::numeric_limits<int64_t>::max()))
      error_at(at, std::string(what) + " tag " + std::to_string(t) +
                       ": tags must be 1 to 2^63 - 1");
    return static_cast<int64_t>(t);
  }

  // Bytes one value takes at least: int, size_t, double.
  size_t w_int() const { return binary_ ? 4 : kToken; }
  size_t w_size() const { return binary_ ? size_t_ : kToken; }
  size_t w_float() const { return binary_ ? 8 : kToken; }

  void end(const std::string& section) {
    size_t at;
    std::string_view t = token(&at);
    if (t != "$End" + section)
      error_at(at, "expected $End" + section + ", found \"" + std::string(t) + "\"");
  }

  void skip_section(const std::string& section) {
    const std::string end = "$End" + section;
    const size_t at = pos_;
    const void* hit = nullptr;
    if (remaining() >= end.size())
      hit = memmem(b_ + pos_, remaining(), end.data(), end.size());
    if (!hit) error_at(at, "no " + end);
    pos_ = static_cast<const unsigned char*>(hit) - b_ + end.size();
  }

  void set_binary(bool big, size_t st) {
    binary_ = true;
    big_ = big;
    size_t_ = st;
  }
  void set_text(bool text) { binary_ = !text && size_t_ != 0; }
  bool big() const { return big_; }

 private:
  const unsigned char* b_;
  size_t n_;
  size_t pos_ = 0;
  bool binary_ = false, big_ = false;
  size_t size_t_ = 0;
};

void mesh_format(Cursor& c) {
  size_t at;
  std::string_view version = c.token(&at);
  if (version != "4.1")
    c.error_at(at, "MSH version " + std::string(version) + ": only 4.1 is supported");
  at = c.pos();
  const int32_t file_type = c.int_();
  const int32_t data_size = c.int_();
  if (data_size != 4 && data_size != 8)
    c.error_at(at, "data size " + std::to_string(data_size) + " is not 4 or 8");
  if (file_type == 1) {
    c.set_binary(false, static_cast<size_t>(data_size));
    c.header_end();
    const size_t at1 = c.pos();
    const unsigned char* one = c.take(4);
    if (one[0] == 1 && one[1] == 0 && one[2] == 0 && one[3] == 0)
      c.set_binary(false, static_cast<size_t>(data_size));
    else if (one[0] == 0 && one[1] == 0 && one[2] == 0 && one[3] == 1)
      c.set_binary(true, static_cast<size_t>(data_size));
    else
      c.error_at(at1, "an endianness marker that is not 1");
  } else if (file_type != 0) {
    c.error_at(at, "file type " + std::to_string(file_type) + " is not 0 or 1");
  }
  c.end("MeshFormat");
}

std::vector<PhysicalName> physical_names(Cursor& c) {
  const bool binary = c.binary();
  c.set_text(true);
  std::vector<PhysicalName> names;
  const size_t n = c.count(3 * kToken + 2);
  for (size_t i = 0; i < n; ++i) {
    PhysicalName p;
    p.dim = c.dim();
    p.tag = c.int_();
    c.skip_whitespace();
    const size_t at = c.pos();
    if (c.remaining() == 0 || *c.take(1) != '"') c.error_at(at, "expected a quoted name");
    std::string name;
    bool closed = false;
    while (c.remaining() > 0) {
      const unsigned char ch = *c.take(1);
      if (ch == '"') {
        closed = true;
        break;
      }
      name.push_back(static_cast<char>(ch));
    }
    if (!closed) c.error_at(at, "unterminated name");
    if (name.find_first_of("\r\n") != std::string::npos) c.error_at(at, "a name with a line break");
    p.name = std::move(name);
    names.push_back(std::move(p));
  }
  c.end("PhysicalNames");
  c.set_text(!binary);
  return names;
}

std::vector<int32_t> int_list(Cursor& c) {
  const size_t n = c.count(c.w_int());
  std::vector<int32_t> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = c.int_();
  return v;
}

void entity_body(Cursor& c, Entity& e, int dim) {
  for (int g = 0; g < 3; ++g) e.min[g] = c.float_();
  for (int g = 0; g < 3; ++g) e.max[g] = dim == 0 ? e.min[g] : c.float_();
  e.physical = int_list(c);
  if (dim > 0) e.boundary = int_list(c);
}

void entities(Cursor& c, File& f) {
  size_t counts[4];
  for (size_t& n : counts) n = c.count(c.w_int() + 3 * c.w_float());
  for (int dim = 0; dim < 4; ++dim) {
    f.entities[dim].clear();
    for (size_t i = 0; i < counts[dim]; ++i) {
      Entity e;
      e.tag = c.int_();
      entity_body(c, e, dim);
      f.entities[dim].push_back(std::move(e));
    }
  }
  f.has_entities = true;
  c.end("Entities");
}

void partitioned_entities(Cursor& c, File& f) {
  f.num_partitions = c.count(0);
  const size_t ghosts = c.count(2 * c.w_int());
  f.ghost_entities.clear();
  for (size_t i = 0; i < ghosts; ++i) {
    const int32_t tag = c.int_();
    const int32_t part = c.int_();
    f.ghost_entities.push_back({tag, part});
  }
  size_t counts[4];
  for (size_t& n : counts) n = c.count(3 * c.w_int() + 3 * c.w_float());
  for (int dim = 0; dim < 4; ++dim) {
    f.partitioned[dim].clear();
    for (size_t i = 0; i < counts[dim]; ++i) {
      PartitionedEntity p;
      p.entity.tag = c.int_();
      p.parent_dim = c.dim();
      p.parent_tag = c.int_();
      p.partitions = int_list(c);
      entity_body(c, p.entity, dim);
      f.partitioned[dim].push_back(std::move(p));
    }
  }
  f.has_partitioned = true;
  c.end("PartitionedEntities");
}

void periodic(Cursor& c, File& f) {
  const size_t n = c.count(3 * c.w_int() + 2 * c.w_size());
  for (size_t i = 0; i < n; ++i) {
    Periodic p;
    p.dim = c.dim();
    p.tag = c.int_();
    p.master_tag = c.int_();
    const size_t na = c.count(c.w_float());
    for (size_t k = 0; k < na; ++k) p.affine.push_back(c.float_());
    const size_t nn = c.count(2 * c.w_size());
    for (size_t k = 0; k < nn; ++k) {
      const int64_t a = c.tag("node");
      const int64_t b = c.tag("node");
      p.nodes.push_back({a, b});
    }
    f.periodic.push_back(std::move(p));
  }
  c.end("Periodic");
}

void nodes(Cursor& c, File& f) {
  const size_t at = c.pos();
  const size_t num_blocks = c.count(3 * c.w_int() + c.w_size());
  const uint64_t num_nodes = c.size();
  (void)c.size();  // minNodeTag
  (void)c.size();  // maxNodeTag
  uint64_t total = 0;
  for (size_t b = 0; b < num_blocks; ++b) {
    NodeBlock blk;
    blk.dim = c.dim();
    blk.entity_tag = c.int_();
    const size_t at_param = c.pos();
    const int32_t param = c.int_();
    if (param != 0 && param != 1)
      c.error_at(at_param, "parametric flag " + std::to_string(param) + " is not 0 or 1");
    const int parametric = param == 1 ? blk.dim : 0;
    const size_t n = c.count(c.w_size() + 3 * c.w_float());
    blk.tags.resize(n);
    c.tags_into(blk.tags.data(), n, "node");
    blk.coords.resize(3 * n);
    if (parametric == 0) {
      c.floats_into(blk.coords.data(), 3 * n);
    } else {
      // Parametric coordinates follow each node's x, y, z; they are read and dropped.
      std::vector<double> row(3 + parametric);
      for (size_t i = 0; i < n; ++i) {
        c.floats_into(row.data(), row.size());
        for (int g = 0; g < 3; ++g) blk.coords[3 * i + g] = row[g];
      }
    }
    total += n;
    f.node_blocks.push_back(std::move(blk));
  }
  if (total != num_nodes)
    c.error_at(at, std::to_string(num_nodes) + " nodes announced, " + std::to_string(total) +
                       " found");
  c.end("Nodes");
}

void elements(Cursor& c, File& f) {
  const size_t at = c.pos();
  const size_t num_blocks = c.count(3 * c.w_int() + c.w_size());
  const uint64_t num_elements = c.size();
  (void)c.size();
  (void)c.size();
  uint64_t total = 0;
  for (size_t b = 0; b < num_blocks; ++b) {
    ElementBlock blk;
    blk.dim = c.dim();
    blk.entity_tag = c.int_();
    const size_t at_type = c.pos();
    blk.element_type = c.int_();
    const ElementType ty = element_type(blk.element_type);
    if (tdim(ty.cell) != blk.dim)
      c.error_at(at_type, "element type " + std::to_string(blk.element_type) +
                              " in a block of dimension " + std::to_string(blk.dim));
    const size_t k = ty.num_nodes;
    const size_t n = c.count((1 + k) * c.w_size());
    blk.tags.resize(n);
    blk.nodes.resize(n * k);
    if (c.binary()) {
      c.tags_split(blk.nodes.data(), blk.tags.data(), n, k, "element or node");
    } else {
      for (size_t i = 0; i < n; ++i) {
        blk.tags[i] = c.tag("element");
        for (size_t j = 0; j < k; ++j) blk.nodes[i * k + j] = c.tag("node");
      }
    }
    total += n;
    f.element_blocks.push_back(std::move(blk));
  }
  if (total != num_elements)
    c.error_at(at, std::to_string(num_elements) + " elements announced, " +
                       std::to_string(total) + " found");
  c.end("Elements");
}

}  // namespace

File parse(const unsigned char* bytes, size_t len) {
  Cursor c(bytes, len);
  size_t at;
  std::string_view first = c.token(&at);
  if (first != "$MeshFormat") c.error_at(at, "not a Gmsh file: no $MeshFormat");
  mesh_format(c);
  File f;
  f.binary = c.binary();
  while (!c.at_end()) {
    std::string_view header = c.token(&at);
    if (header.empty() || header[0] != '$')
      c.error_at(at, "expected a section, found \"" + std::string(header) + "\"");
    const std::string section(header.substr(1));
    c.header_end();
    if (section == "PhysicalNames") {
      std::vector<PhysicalName> n = physical_names(c);
      f.names.insert(f.names.end(), n.begin(), n.end());
    } else if (section == "Entities") {
      entities(c, f);
    } else if (section == "PartitionedEntities") {
      partitioned_entities(c, f);
    } else if (section == "Periodic") {
      periodic(c, f);
    } else if (section == "Nodes") {
      nodes(c, f);
    } else if (section == "Elements") {
      elements(c, f);
    } else if (section == "MeshFormat") {
      c.error_at(at, "a second $MeshFormat");
    } else if (section.compare(0, 3, "End") == 0) {
      c.error_at(at, std::string(header) + " without its section");
    } else {
      c.skip_section(section);
    }
  }
  (void)kUnsupportedElementType;
  return f;
}

}  // namespace tg::gmsh
