#include "yaml.hpp"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace cc4s_amd {
namespace yaml {

NodePtr Node::scalar(const std::string &v, bool quoted) {
  auto n = std::make_shared<Node>();
  n->kind = Kind::Scalar;
  n->value = v;
  n->quoted = quoted;
  return n;
}

NodePtr Node::map() { return std::make_shared<Node>(); }

NodePtr Node::get(const std::string &key) const {
  if (kind != Kind::Map) return nullptr;
  for (const auto &e : entries)
    if (e.first == key) return e.second;
  return nullptr;
}

void Node::set(const std::string &key, NodePtr v) {
  for (auto &e : entries)
    if (e.first == key) {
      e.second = std::move(v);
      return;
    }
  entries.emplace_back(key, std::move(v));
}

void Node::set(const std::string &key, double v) { set(key, scalar(format_real(v))); }
void Node::set(const std::string &key, int64_t v) { set(key, scalar(std::to_string(v))); }

namespace {

std::runtime_error error_at(int line, const std::string &what) {
  return std::runtime_error("line " + std::to_string(line) + ": " + what);
}

}  // namespace

std::string Node::text() const {
  if (kind != Kind::Scalar) throw error_at(line, "expecting a scalar");
  return value;
}

double Node::real() const {
  const std::string t = text();
  char *end = nullptr;
  errno = 0;
  const double v = std::strtod(t.c_str(), &end);
  if (end == t.c_str() || *end != '\0' || errno == ERANGE)
    throw error_at(line, "expecting a real number, found '" + t + "'");
  return v;
}

int64_t Node::integer() const {
  const std::string t = text();
  char *end = nullptr;
  errno = 0;
  const long long v = std::strtoll(t.c_str(), &end, 10);
  if (end == t.c_str() || *end != '\0' || errno == ERANGE)
    throw error_at(line, "expecting an integer, found '" + t + "'");
  return v;
}

bool Node::boolean() const {
  const std::string t = text();
  if (t == "1" || t == "true" || t == "True" || t == "TRUE" || t == "yes") return true;
  if (t == "0" || t == "false" || t == "False" || t == "FALSE" || t == "no") return false;
  throw error_at(line, "expecting a boolean, found '" + t + "'");
}

std::string get_text(const NodePtr &m, const std::string &key, const std::string &dflt) {
  auto n = m ? m->get(key) : nullptr;
  return n ? n->text() : dflt;
}
double get_real(const NodePtr &m, const std::string &key, double dflt) {
  auto n = m ? m->get(key) : nullptr;
  return n ? n->real() : dflt;
}
int64_t get_integer(const NodePtr &m, const std::string &key, int64_t dflt) {
  auto n = m ? m->get(key) : nullptr;
  return n ? n->integer() : dflt;
}
bool get_boolean(const NodePtr &m, const std::string &key, bool dflt) {
  auto n = m ? m->get(key) : nullptr;
  return n ? n->boolean() : dflt;
}

// --- parser ------------------------------------------------------------------

namespace {

struct Line {
  int number;       // 1-based
  int indent;       // leading spaces
  std::string text; // without indentation and comment
};

std::string rtrim(const std::string &s) {
  size_t e = s.find_last_not_of(" \t\r");
  return e == std::string::npos ? std::string() : s.substr(0, e + 1);
}

std::string trim(const std::string &s) {
  size_t b = s.find_first_not_of(" \t\r");
  if (b == std::string::npos) return std::string();
  return rtrim(s.substr(b));
}

// removes a comment: '#' at the start or after whitespace, outside quotes
std::string strip_comment(const std::string &s) {
  char quote = 0;
  for (size_t i = 0; i < s.size(); ++i) {
    const char c = s[i];
    if (quote) {
      if (c == quote) quote = 0;
    } else if (c == '"' || c == '\'') {
      quote = c;
    } else if (c == '#' && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t')) {
      return s.substr(0, i);
    }
  }
  return s;
}

class Parser {
 public:
  Parser(const std::string &text, const std::string &source) : source_(source) {
    std::istringstream in(text);
    std::string raw;
    int number = 0;
    while (std::getline(in, raw)) {
      ++number;
      if (raw.find('\t') != std::string::npos &&
          raw.find_first_not_of(" \t") != std::string::npos &&
          raw.find('\t') < raw.find_first_not_of(" \t"))
        throw err(number, "tab in indentation");
      std::string s = rtrim(strip_comment(raw));
      const size_t b = s.find_first_not_of(' ');
      if (b == std::string::npos) continue;
      const std::string body = s.substr(b);
      if (body == "---" || body == "...") continue;
      lines_.push_back({number, static_cast<int>(b), body});
    }
  }

  NodePtr parse() {
    if (lines_.empty()) return Node::map();
    NodePtr n = block(lines_[0].indent);
    if (pos_ != lines_.size()) throw err(lines_[pos_].number, "unexpected indentation");
    return n;
  }

 private:
  std::runtime_error err(int line, const std::string &what) const {
    return std::runtime_error(source_ + ":" + std::to_string(line) + ": " + what);
  }

  static bool is_item(const std::string &t) { return t == "-" || t.rfind("- ", 0) == 0; }

  // position of the ':' that separates key and value (outside quotes and
  // brackets, followed by a space or the end), or npos
  static size_t key_colon(const std::string &t) {
    char quote = 0;
    int depth = 0;
    for (size_t i = 0; i < t.size(); ++i) {
      const char c = t[i];
      if (quote) {
        if (c == quote) quote = 0;
        continue;
      }
      if (c == '"' || c == '\'') {
        if (i == 0) quote = c;
        continue;
      }
      if (c == '{' || c == '[') ++depth;
      if (c == '}' || c == ']') --depth;
      if (c == ':' && depth == 0 && (i + 1 == t.size() || t[i + 1] == ' ')) return i;
    }
    return std::string::npos;
  }

  // the value of "key:" on the following, deeper lines: a block sequence, a
  // block map, or one line holding a flow collection or a scalar (the
  // reference's operation files write "out:" and then "{}")
  NodePtr block(int indent) {
    if (pos_ >= lines_.size()) return Node::map();
    const Line &l = lines_[pos_];
    if (is_item(l.text)) return sequence(indent);
    if (l.text[0] == '{' || l.text[0] == '[' || key_colon(l.text) == std::string::npos) {
      ++pos_;
      NodePtr v = flow_or_scalar(l.text, l.number);
      if (pos_ < lines_.size() && lines_[pos_].indent >= indent)
        throw err(lines_[pos_].number, "unexpected indentation");
      return v;
    }
    return mapping(indent);
  }

  NodePtr sequence(int indent) {
    auto m = Node::map();
    m->line = lines_[pos_].number;
    size_t index = 0;
    while (pos_ < lines_.size() && lines_[pos_].indent == indent && is_item(lines_[pos_].text)) {
      Line &l = lines_[pos_];
      const std::string rest = l.text == "-" ? std::string() : trim(l.text.substr(2));
      NodePtr item;
      if (rest.empty()) {
        ++pos_;
        item = (pos_ < lines_.size() && lines_[pos_].indent > indent) ? block(lines_[pos_].indent)
                                                                     : Node::scalar("");
      } else if (key_colon(rest) != std::string::npos && rest[0] != '{' && rest[0] != '[') {
        // "- key: value": a map whose first entry sits on the item line
        const int inner = indent + static_cast<int>(l.text.size() - rest.size());
        l.indent = inner;
        l.text = rest;
        item = mapping(inner);
      } else {
        ++pos_;
        item = flow_or_scalar(rest, l.number);
      }
      m->set(std::to_string(index++), item);
    }
    return m;
  }

  NodePtr mapping(int indent) {
    auto m = Node::map();
    m->line = lines_[pos_].number;
    while (pos_ < lines_.size() && lines_[pos_].indent == indent) {
      const Line l = lines_[pos_];
      if (is_item(l.text)) throw err(l.number, "sequence item inside a map");
      const size_t c = key_colon(l.text);
      if (c == std::string::npos) throw err(l.number, "expecting 'key: value'");
      std::string key = trim(l.text.substr(0, c));
      if (key.size() >= 2 && (key[0] == '"' || key[0] == '\'') && key.back() == key[0])
        key = key.substr(1, key.size() - 2);
      const std::string rest = trim(l.text.substr(c + 1));
      ++pos_;
      NodePtr value;
      if (!rest.empty()) {
        value = flow_or_scalar(rest, l.number);
      } else if (pos_ < lines_.size() && lines_[pos_].indent > indent) {
        value = block(lines_[pos_].indent);
      } else if (pos_ < lines_.size() && lines_[pos_].indent == indent &&
                 is_item(lines_[pos_].text)) {
        value = sequence(indent);  // "key:" followed by "- item" at the same indent
      } else {
        value = Node::scalar("");
      }
      if (m->has(key)) throw err(l.number, "duplicate key '" + key + "'");
      m->set(key, value);
    }
    if (pos_ < lines_.size() && lines_[pos_].indent > indent)
      throw err(lines_[pos_].number, "unexpected indentation");
    return m;
  }

  NodePtr flow_or_scalar(const std::string &t, int line) {
    size_t p = 0;
    NodePtr n = flow(t, p, line);
    while (p < t.size() && t[p] == ' ') ++p;
    if (p != t.size()) throw err(line, "trailing text after value");
    return n;
  }

  NodePtr flow(const std::string &t, size_t &p, int line) {
    while (p < t.size() && t[p] == ' ') ++p;
    if (p >= t.size()) return Node::scalar("");
    if (t[p] == '{' || t[p] == '[') {
      const char close = t[p] == '{' ? '}' : ']';
      const bool is_map = t[p] == '{';
      ++p;
      auto m = Node::map();
      m->line = line;
      size_t index = 0;
      while (true) {
        while (p < t.size() && t[p] == ' ') ++p;
        if (p >= t.size()) throw err(line, "unterminated flow collection");
        if (t[p] == close) {
          ++p;
          break;
        }
        if (is_map) {
          // key up to ':'
          size_t k0 = p;
          std::string key;
          if (t[p] == '"' || t[p] == '\'') {
            key = quoted(t, p, line);
          } else {
            while (p < t.size() && t[p] != ':' && t[p] != ',' && t[p] != close) ++p;
            key = trim(t.substr(k0, p - k0));
          }
          while (p < t.size() && t[p] == ' ') ++p;
          if (p >= t.size() || t[p] != ':') throw err(line, "expecting ':' in flow map");
          ++p;
          m->set(key, flow(t, p, line));
        } else {
          m->set(std::to_string(index++), flow(t, p, line));
        }
        while (p < t.size() && t[p] == ' ') ++p;
        if (p < t.size() && t[p] == ',') ++p;
      }
      return m;
    }
    NodePtr s;
    if (t[p] == '"' || t[p] == '\'') {
      s = Node::scalar(quoted(t, p, line), true);
    } else {
      size_t b = p;
      // plain scalar: in flow context it ends at ',', '}' or ']'
      int depth = 0;
      while (p < t.size()) {
        const char c = t[p];
        if (c == '(') ++depth;
        if (c == ')') --depth;
        if (depth == 0 && (c == ',' || c == '}' || c == ']')) break;
        ++p;
      }
      s = Node::scalar(trim(t.substr(b, p - b)));
    }
    s->line = line;
    return s;
  }

  std::string quoted(const std::string &t, size_t &p, int line) {
    const char q = t[p++];
    std::string out;
    while (p < t.size()) {
      const char c = t[p++];
      if (q == '\'' && c == '\'') {
        if (p < t.size() && t[p] == '\'') {
          out += '\'';
          ++p;
          continue;
        }
        return out;
      }
      if (q == '"' && c == '"') return out;
      if (q == '"' && c == '\\' && p < t.size()) {
        const char e = t[p++];
        out += e == 'n' ? '\n' : e == 't' ? '\t' : e;
        continue;
      }
      out += c;
    }
    throw err(line, "unterminated quoted scalar");
  }

  std::string source_;
  std::vector<Line> lines_;
  size_t pos_ = 0;
};

bool needs_quotes(const std::string &v) {
  if (v.empty()) return true;
  if (v.find_first_of(":#{}[],&*!|>'\"%@`") != std::string::npos) return true;
  if (v.front() == ' ' || v.back() == ' ' || v.front() == '-') {
    // negative numbers stay plain
    char *end = nullptr;
    std::strtod(v.c_str(), &end);
    return !(end && *end == '\0');
  }
  return false;
}

void emit_node(std::ostringstream &out, const NodePtr &n, int indent) {
  const std::string pad(indent, ' ');
  for (const auto &e : n->entries) {
    out << pad << e.first << ":";
    const NodePtr &v = e.second;
    if (!v) {
      out << " ~\n";
    } else if (v->is_scalar()) {
      if (needs_quotes(v->value)) {
        std::string q;
        for (char c : v->value) q += (c == '"' || c == '\\') ? std::string("\\") + c : std::string(1, c);
        out << " \"" << q << "\"\n";
      } else {
        out << " " << v->value << "\n";
      }
    } else if (v->entries.empty()) {
      out << " {}\n";
    } else {
      out << "\n";
      emit_node(out, v, indent + 2);
    }
  }
}

}  // namespace

NodePtr parse(const std::string &text, const std::string &source_name) {
  return Parser(text, source_name).parse();
}

NodePtr parse_file(const std::string &path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("Failed to open file \"" + path + "\"");
  std::stringstream ss;
  ss << in.rdbuf();
  return parse(ss.str(), path);
}

std::string emit(const NodePtr &node) {
  std::ostringstream out;
  if (node && node->is_map()) {
    if (node->entries.empty()) return "{}\n";
    emit_node(out, node, 0);
  } else if (node) {
    out << node->value << "\n";
  }
  return out.str();
}

void emit_file(const NodePtr &node, const std::string &path) {
  std::ofstream out(path);
  if (!out) throw std::runtime_error("Failed to open file \"" + path + "\"");
  out << emit(node);
}

std::string format_real(double v) {
  char buf[40];
  for (int digits = 15; digits <= 17; ++digits) {
    std::snprintf(buf, sizeof(buf), "%.*g", digits, v);
    if (std::strtod(buf, nullptr) == v) break;
  }
  std::string s(buf);
  // keep a decimal point so that the value reads back as a real
  if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
  return s;
}

}  // namespace yaml
}  // namespace cc4s_amd
