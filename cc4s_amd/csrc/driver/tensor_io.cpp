#include "tensor_io.hpp"

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace cc4s_amd {

int64_t HostTensor::elements() const {
  int64_t n = 1;
  for (auto l : lens) n *= l;
  return n;
}

namespace {

std::string directory_prefix(const std::string &path) {
  const auto slash = path.rfind('/');
  return slash == std::string::npos ? std::string() : path.substr(0, slash + 1);
}

// base name without directory and extension (Reader.hpp / Writer.hpp)
std::string base_name(const std::string &path) {
  const auto dot = path.rfind('.');
  const auto slash = path.rfind('/');
  if (dot == std::string::npos || (slash != std::string::npos && dot < slash))
    throw std::runtime_error("'fileName' must contain an extension e.g. '.yaml'");
  const size_t b = slash == std::string::npos ? 0 : slash + 1;
  return path.substr(b, dot - b);
}

void read_binary(const std::string &file, HostTensor &t) {
  std::ifstream in(file, std::ios::binary | std::ios::ate);
  if (!in) throw std::runtime_error("Failed to open file '" + file + "'");
  const std::streamoff bytes = in.tellg();
  const int64_t want = t.elements() * (t.complex ? 16 : 8);
  if (bytes != want)
    throw std::runtime_error("binary file '" + file + "' holds " + std::to_string(bytes) +
                             " bytes, the header needs " + std::to_string(want));
  in.seekg(0);
  t.data.resize(static_cast<size_t>(t.elements()) * (t.complex ? 2 : 1));
  in.read(reinterpret_cast<char *>(t.data.data()), want);
  if (!in) throw std::runtime_error("failed reading '" + file + "'");
}

// numbers separated by white space; complex numbers as "(re,im)" or a plain
// real number (imaginary part 0)
void read_text(const std::string &file, HostTensor &t) {
  std::ifstream in(file);
  if (!in) throw std::runtime_error("Failed to open file \"" + file + "\"");
  std::stringstream ss;
  ss << in.rdbuf();
  const std::string s = ss.str();
  const int64_t n = t.elements();
  t.data.assign(static_cast<size_t>(n) * (t.complex ? 2 : 1), 0.0);
  const char *p = s.c_str();
  auto number = [&](const char *&q) {
    char *end = nullptr;
    const double v = std::strtod(q, &end);
    if (end == q) throw std::runtime_error("expecting a number in '" + file + "'");
    q = end;
    return v;
  };
  auto skip = [&](const char *&q) {
    while (*q && (std::isspace(static_cast<unsigned char>(*q)))) ++q;
    // comment lines ("# dry-run ...")
    while (*q == '#') {
      while (*q && *q != '\n') ++q;
      while (*q && std::isspace(static_cast<unsigned char>(*q))) ++q;
    }
  };
  for (int64_t e = 0; e < n; ++e) {
    skip(p);
    if (!*p)
      throw std::runtime_error("text file '" + file + "' ends after " + std::to_string(e) +
                               " of " + std::to_string(n) + " elements");
    if (*p == '(') {
      ++p;
      const double re = number(p);
      skip(p);
      if (*p != ',') throw std::runtime_error("expecting ',' in '" + file + "'");
      ++p;
      const double im = number(p);
      skip(p);
      if (*p != ')') throw std::runtime_error("expecting ')' in '" + file + "'");
      ++p;
      if (!t.complex && im != 0.0)
        throw std::runtime_error("complex element in the real tensor '" + file + "'");
      if (t.complex) {
        t.data[2 * e] = re;
        t.data[2 * e + 1] = im;
      } else {
        t.data[e] = re;
      }
    } else {
      const double v = number(p);
      if (t.complex)
        t.data[2 * e] = v;
      else
        t.data[e] = v;
    }
  }
}

void write_elements(const HostTensor &t, const std::string &file, bool binary) {
  if (binary) {
    std::ofstream out(file, std::ios::binary);
    if (!out) throw std::runtime_error("Failed to open file '" + file + "'");
    out.write(reinterpret_cast<const char *>(t.data.data()),
              static_cast<std::streamsize>(t.data.size() * sizeof(double)));
    return;
  }
  std::FILE *f = std::fopen(file.c_str(), "w");
  if (!f) throw std::runtime_error("Failed to open file \"" + file + "\"");
  const int64_t n = t.elements();
  for (int64_t e = 0; e < n; ++e) {
    if (t.complex)
      std::fprintf(f, "(%.17g,%.17g)\n", t.data[2 * e], t.data[2 * e + 1]);
    else
      std::fprintf(f, "%.17g\n", t.data[e]);
  }
  std::fclose(f);
}

}  // namespace

HostTensor read_tensor(const yaml::NodePtr &header, const std::string &node_path) {
  HostTensor t;
  const int64_t version = yaml::get_integer(header, "version", -1);
  if (version != kTensorVersion)
    throw std::runtime_error("Found and expected serialization versions mismatch. Found " +
                             std::to_string(version) + ", expecting " +
                             std::to_string(kTensorVersion));
  const std::string scalar = yaml::get_text(header, "scalarType", "");
  if (scalar == "Complex64")
    t.complex = true;
  else if (scalar != "Real64")
    throw std::runtime_error("Unsupported scalarType \"" + scalar + "\"");
  const auto dims = header->get("dimensions");
  if (!dims || !dims->is_map()) throw std::runtime_error("tensor header without dimensions");
  for (const auto &d : dims->entries) {
    t.lens.push_back(yaml::get_integer(d.second, "length", -1));
    if (t.lens.back() < 0) throw std::runtime_error("dimension without length");
    t.dimension_types.push_back(yaml::get_text(d.second, "type", ""));
  }
  if (header->has("nonZeroCondition"))
    throw std::runtime_error("block-sparse tensors (nonZeroCondition) are not supported");
  t.unit = yaml::get_real(header, "unit", 1.0);
  t.meta_data = header->get("metaData");
  const std::string elements = yaml::get_text(header->get("elements"), "type", "TextFile");
  const std::string file = node_path + ".elements";
  if (elements == "IeeeBinaryFile")
    read_binary(file, t);
  else
    read_text(file, t);
  return t;
}

yaml::NodePtr write_tensor(const HostTensor &t, const std::string &node_path, bool binary) {
  auto h = yaml::Node::map();
  h->set("type", "Tensor");
  auto dims = yaml::Node::map();
  for (size_t d = 0; d < t.lens.size(); ++d) {
    auto dim = yaml::Node::map();
    dim->set("length", t.lens[d]);
    if (d < t.dimension_types.size() && !t.dimension_types[d].empty())
      dim->set("type", t.dimension_types[d]);
    dims->set(std::to_string(d), dim);
  }
  h->set("dimensions", dims);
  h->set("scalarType", t.scalar_type());
  h->set("version", kTensorVersion);
  auto el = yaml::Node::map();
  el->set("type", binary ? "IeeeBinaryFile" : "TextFile");
  h->set("elements", el);
  h->set("unit", t.unit);
  h->set("metaData", t.meta_data ? t.meta_data : yaml::Node::map());
  write_elements(t, node_path + ".elements", binary);
  return h;
}

ReadItem read_file(const std::string &path_file_name) {
  const std::string dir = directory_prefix(path_file_name);
  const std::string base = dir + base_name(path_file_name);
  auto root = yaml::parse_file(path_file_name);
  ReadItem item;
  const std::string type = root->is_map() ? yaml::get_text(root, "type", "") : "";
  if (type == "Tensor") {
    item.kind = ReadItem::Kind::Tensor;
    item.tensor = read_tensor(root, base);
  } else if (type == "TensorSet") {
    item.kind = ReadItem::Kind::TensorSet;
    const auto comps = root->get("components");
    if (!comps || !comps->is_map()) throw std::runtime_error("TensorSet without components");
    for (const auto &c : comps->entries)
      item.set[c.first] = read_tensor(c.second, base + ".components." + c.first);
  } else {
    item.data = root;
  }
  return item;
}

yaml::NodePtr write_tensor_file(const HostTensor &t, const std::string &file_name, bool binary) {
  const std::string base = directory_prefix(file_name) + base_name(file_name);
  auto h = write_tensor(t, base, binary);
  yaml::emit_file(h, file_name);
  return h;
}

yaml::NodePtr write_tensor_set_file(const HostTensorSet &s, const std::string &file_name,
                                    bool binary) {
  const std::string base = directory_prefix(file_name) + base_name(file_name);
  auto root = yaml::Node::map();
  auto comps = yaml::Node::map();
  for (const auto &c : s) comps->set(c.first, write_tensor(c.second, base + ".components." + c.first, binary));
  root->set("components", comps);
  root->set("type", "TensorSet");
  yaml::emit_file(root, file_name);
  return root;
}

}  // namespace cc4s_amd
