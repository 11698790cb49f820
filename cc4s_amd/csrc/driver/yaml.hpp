// A YAML subset for cc4s operation files, tensor headers and the output
// report.
//
// cc4s keeps every YAML document as a tree of maps and scalars: sequences are
// read as maps with the keys "0", "1", ... (Parser.hpp, parseSequence) and
// written back as maps (Emitter.hpp, emitMap).  This reader accepts block
// maps, block sequences ("- "), flow maps ("{a: 1}", "{}"), flow sequences
// ("[1, 2]"), plain, single- and double-quoted scalars and comments; that is
// what cc4s writes and what its operation files use.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace cc4s_amd {
namespace yaml {

struct Node;
using NodePtr = std::shared_ptr<Node>;

struct Node {
  enum class Kind { Scalar, Map };
  Kind kind = Kind::Map;
  std::string value;  // scalar text (quotes removed)
  bool quoted = false;
  std::vector<std::pair<std::string, NodePtr>> entries;  // map, in order
  int line = 0;                                          // source line (1-based)

  static NodePtr scalar(const std::string &v, bool quoted = false);
  static NodePtr map();

  bool is_map() const { return kind == Kind::Map; }
  bool is_scalar() const { return kind == Kind::Scalar; }
  size_t size() const { return entries.size(); }

  // map access; null when absent
  NodePtr get(const std::string &key) const;
  NodePtr get(size_t index) const { return get(std::to_string(index)); }
  bool has(const std::string &key) const { return get(key) != nullptr; }
  // inserts or replaces
  void set(const std::string &key, NodePtr v);
  void set(const std::string &key, const std::string &v) { set(key, scalar(v)); }
  void set(const std::string &key, const char *v) { set(key, scalar(v)); }
  void set(const std::string &key, double v);
  void set(const std::string &key, int64_t v);

  // typed scalar values; throw with the source line on mismatch
  std::string text() const;
  double real() const;
  int64_t integer() const;
  bool boolean() const;  // 0/1, true/false
};

// typed lookups with defaults (the reference's getValue<T>(key, default))
std::string get_text(const NodePtr &m, const std::string &key, const std::string &dflt);
double get_real(const NodePtr &m, const std::string &key, double dflt);
int64_t get_integer(const NodePtr &m, const std::string &key, int64_t dflt);
bool get_boolean(const NodePtr &m, const std::string &key, bool dflt);

NodePtr parse(const std::string &text, const std::string &source_name);
NodePtr parse_file(const std::string &path);
std::string emit(const NodePtr &node);
void emit_file(const NodePtr &node, const std::string &path);
// shortest decimal that reads back to the same double (17 significant digits)
std::string format_real(double v);

}  // namespace yaml
}  // namespace cc4s_amd
