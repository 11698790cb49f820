// cc4s tensor files: a YAML header (TensorIo.cxx:114-158 writes it,
// :160-246 reads it) and the elements beside it, either raw little-endian
// numbers ("IeeeBinaryFile", the dense file of CtfMachineTensor.hpp:182-205)
// or one number per line ("TextFile", 17 significant digits, complex numbers
// as "(re,im)", TensorIo.cxx:248-300).  Elements are dense with the first
// index fastest (reference/binary_tensor_file_format.txt:86-95).
#pragma once

#include "yaml.hpp"

#include <map>
#include <string>
#include <vector>

namespace cc4s_amd {

struct HostTensor {
  bool complex = false;                  // scalarType Complex64 / Real64
  std::vector<int64_t> lens;
  std::vector<std::string> dimension_types;  // "State", "AuxiliaryField", ... or ""
  double unit = 1.0;
  yaml::NodePtr meta_data;               // may be null
  std::vector<double> data;              // interleaved (re, im) when complex

  int64_t elements() const;
  const char *scalar_type() const { return complex ? "Complex64" : "Real64"; }
};

using HostTensorSet = std::map<std::string, HostTensor>;

// serialization version of TensorIo.cxx:27
constexpr int64_t kTensorVersion = 100;

// header -> tensor, elements from `node_path` + ".elements"
HostTensor read_tensor(const yaml::NodePtr &header, const std::string &node_path);
// writes `node_path` + ".elements" and returns the header node
yaml::NodePtr write_tensor(const HostTensor &t, const std::string &node_path, bool binary);

// A file read as the reference's Reader does (Reader.hpp): the YAML tree,
// with every map of type Tensor / TensorSet replaced by its data.
struct ReadItem {
  enum class Kind { Data, Tensor, TensorSet } kind = Kind::Data;
  yaml::NodePtr data;
  HostTensor tensor;
  HostTensorSet set;
};
ReadItem read_file(const std::string &path_file_name);

// Writer.hpp for a tensor or a tensor set: `file_name` (with extension)
// receives the header, the elements go beside it as base.elements or
// base.components.<key>.elements
yaml::NodePtr write_tensor_file(const HostTensor &t, const std::string &file_name, bool binary);
yaml::NodePtr write_tensor_set_file(const HostTensorSet &s, const std::string &file_name,
                                    bool binary);

}  // namespace cc4s_amd
