// Device buffers and strided tensor views.
//
// Every tensor of the reference is dense with the first index fastest
// (reference/binary_tensor_file_format.txt:86-95, tcc/Tensor.hpp:114-121);
// a View names its dimensions with one character each (the reference's
// Einstein index strings, e.g. "abij") and carries per-dimension strides, so
// slices (SliceOperator.cxx:112-118), transposes and the conjugate transposes
// of Ccsd.cxx:431-436 are views of the same memory, never copies.
#pragma once

#include "common.hpp"

#include <string>
#include <utility>
#include <cstdlib>
#include <vector>

namespace cc4s_amd {

class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t bytes) { allocate(bytes); }
  ~DeviceBuffer() { release(); }
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  DeviceBuffer(DeviceBuffer &&o) noexcept : p_(o.p_), bytes_(o.bytes_) {
    o.p_ = nullptr;
    o.bytes_ = 0;
  }
  DeviceBuffer &operator=(DeviceBuffer &&o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      bytes_ = o.bytes_;
      o.p_ = nullptr;
      o.bytes_ = 0;
    }
    return *this;
  }
  void allocate(size_t bytes) {
    release();
    if (bytes == 0) return;
    CC4S_HIP(hipMalloc(&p_, bytes));
    bytes_ = bytes;
  }
  // grow-only reallocation, contents are not preserved
  void reserve(size_t bytes) {
    if (bytes > bytes_) allocate(bytes);
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    bytes_ = 0;
  }
  void *get() const { return p_; }
  template <typename T>
  T *as() const {
    return static_cast<T *>(p_);
  }
  size_t bytes() const { return bytes_; }

 private:
  void *p_ = nullptr;
  size_t bytes_ = 0;
};

// host memory for data that does not fit in HBM: pinned (hipHostMalloc)
// where the system grants it, else pageable (the runtime stages the copies;
// slower, but the amount of pinned memory a process may hold is limited)
class HostBuffer {
 public:
  HostBuffer() = default;
  explicit HostBuffer(size_t bytes, bool try_pinned = true) : bytes_(bytes) {
    if (!bytes) return;
    if (try_pinned && hipHostMalloc(&p_, bytes, hipHostMallocDefault) == hipSuccess) {
      pinned_ = true;
      return;
    }
    (void)hipGetLastError();  // the failed pinning is not an error of later calls
    p_ = std::aligned_alloc(4096, (bytes + 4095) / 4096 * 4096);
    if (!p_) throw Error("out of host memory");
    pinned_ = false;
  }
  ~HostBuffer() { release(); }
  HostBuffer(const HostBuffer &) = delete;
  HostBuffer &operator=(const HostBuffer &) = delete;
  HostBuffer(HostBuffer &&o) noexcept : p_(o.p_), bytes_(o.bytes_), pinned_(o.pinned_) {
    o.p_ = nullptr;
    o.bytes_ = 0;
  }
  HostBuffer &operator=(HostBuffer &&o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      bytes_ = o.bytes_;
      pinned_ = o.pinned_;
      o.p_ = nullptr;
      o.bytes_ = 0;
    }
    return *this;
  }
  void *get() const { return p_; }
  size_t bytes() const { return bytes_; }
  bool pinned() const { return pinned_; }

 private:
  void release() {
    if (p_) {
      if (pinned_)
        (void)hipHostFree(p_);
      else
        std::free(p_);
    }
    p_ = nullptr;
  }
  void *p_ = nullptr;
  size_t bytes_ = 0;
  bool pinned_ = false;
};

struct View {
  void *ptr = nullptr;           // element (0,...,0)
  Field field = Field::Complex;  // element = double or (re, im) double pair
  std::string idx;               // one index character per dimension
  std::vector<int64_t> lens;
  std::vector<int64_t> strides;  // in elements of the field
  bool conj = false;

  int rank() const { return static_cast<int>(idx.size()); }
  int64_t size() const {
    int64_t s = 1;
    for (auto l : lens) s *= l;
    return s;
  }
  int find(char c) const {
    auto p = idx.find(c);
    return p == std::string::npos ? -1 : static_cast<int>(p);
  }
  int64_t len(char c) const { return lens[find(c)]; }
  int64_t stride(char c) const { return strides[find(c)]; }

  // same memory, dimensions renamed (no data movement)
  View as(const std::string &names) const {
    CC4S_ASSERT(names.size() == idx.size(), "index rename rank mismatch");
    View v = *this;
    v.idx = names;
    return v;
  }
  View conjugated() const {
    View v = *this;
    v.conj = !v.conj;
    return v;
  }
  // dimension d restricted to [begin, end)
  View sliced(int d, int64_t begin, int64_t end) const {
    View v = *this;
    size_t es = scalar_bytes(field);
    v.ptr = static_cast<char *>(ptr) + begin * strides[d] * es;
    v.lens[d] = end - begin;
    return v;
  }
  // reorder dimensions: new dimension k is old dimension order[k]
  View transposed(const std::vector<int> &order) const {
    View v = *this;
    for (size_t k = 0; k < order.size(); ++k) {
      v.idx[k] = idx[order[k]];
      v.lens[k] = lens[order[k]];
      v.strides[k] = strides[order[k]];
    }
    return v;
  }
  bool dense() const {
    int64_t s = 1;
    for (int d = 0; d < rank(); ++d) {
      if (strides[d] != s) return false;
      s *= lens[d];
    }
    return true;
  }
};

// dense first-index-fastest view over `ptr`
inline View dense_view(void *ptr, Field f, const std::string &idx,
                       const std::vector<int64_t> &lens) {
  View v;
  v.ptr = ptr;
  v.field = f;
  v.idx = idx;
  v.lens = lens;
  v.strides.resize(lens.size());
  int64_t s = 1;
  for (size_t d = 0; d < lens.size(); ++d) {
    v.strides[d] = s;
    s *= lens[d];
  }
  return v;
}

// A dense device tensor owning its buffer.
struct Tensor {
  DeviceBuffer buf;
  View v;
  Tensor() = default;
  Tensor(Field f, const std::string &idx, const std::vector<int64_t> &lens) {
    reset(f, idx, lens);
  }
  void reset(Field f, const std::string &idx, const std::vector<int64_t> &lens) {
    int64_t n = 1;
    for (auto l : lens) n *= l;
    buf.allocate(static_cast<size_t>(n) * scalar_bytes(f));
    v = dense_view(buf.get(), f, idx, lens);
  }
  // view with renamed indices, the way the reference writes (*T)["cdij"]
  View operator()(const std::string &names) const { return v.as(names); }
  int64_t size() const { return v.size(); }
  size_t bytes() const { return buf.bytes(); }
};

}  // namespace cc4s_amd
