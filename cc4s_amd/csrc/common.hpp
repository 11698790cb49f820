// Shared definitions for the cc4s CCSD library on MI355X (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace cc4s_amd {

// Errors carry a message like the reference's Exception text; the C ABI
// turns them into a status code plus cc4s_ccsd_last_error().
struct Error : std::runtime_error {
  explicit Error(const std::string &m) : std::runtime_error(m) {}
};

#define CC4S_HIP(call)                                                        \
  do {                                                                        \
    hipError_t err_ = (call);                                                 \
    if (err_ != hipSuccess)                                                   \
      throw ::cc4s_amd::Error(std::string("HIP error '") +                    \
                              hipGetErrorString(err_) + "' at " __FILE__ ":" + \
                              std::to_string(__LINE__) + " in " #call);      \
  } while (0)

#define CC4S_ASSERT(cond, msg)                                                \
  do {                                                                        \
    if (!(cond)) throw ::cc4s_amd::Error(msg);                                \
  } while (0)

// Scalar field of a run: the reference instantiates Ccsd for Real<64> and
// Complex<64> (coupledcluster/Ccsd.cxx:635-639).
enum class Field : int { Real = 0, Complex = 1 };

inline size_t scalar_bytes(Field f) { return f == Field::Complex ? 16 : 8; }

// Sum and maximum over the 64 lanes of a wavefront by butterfly exchanges
// across lanes (lane l combines with lane l^1, l^2, ..., l^32): no LDS, no
// barrier; every lane ends with the same bits and the order is fixed, so
// the result is reproducible.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) v += __shfl_xor(v, m, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) v = fmax(v, __shfl_xor(v, m, 64));
  return v;
}

// Sets the message cc4s_ccsd_last_error(NULL) returns (calls without a
// context).
void set_last_error(const std::string &message);

}  // namespace cc4s_amd
