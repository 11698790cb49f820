// Allocation-free energy entry point: CoupledCluster::getEnergy
// (CoupledCluster.cxx:194-253) of device-resident amplitudes against
// device-resident V_abij, for callers that cannot allocate device memory.
//
// The reduction partials of the energy kernel live in `scratch`, a
// file-scope device array of kScratchBytes that every GPU the library runs
// on holds until the process ends.  Host code never passes the array's name
// to a kernel or a copy (on the host it is a stand-in, not a device
// address): on the first call on a device, with that device current, its
// device address is taken with hipGetSymbolAddress and kept per device.
//
// Every call on one GPU shares that GPU's array.  The energy kernel writes
// every partial it later reads, so no clearing is needed; calls that overlap
// on one GPU (two host threads, or two streams) race on the partials.
#include "../../include/cc4s_ccsd.h"

#include "elementwise.hpp"

#include <mutex>
#include <string>

namespace {

// 64 KiB: the energy reduction needs (2 * 1024 + 2) double2 = 32,800 bytes,
// independent of No and Nv (fixed grid of 1024 blocks)
constexpr size_t kScratchBytes = 64 * 1024;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

// device address of this device's `scratch`, taken once per device
void *scratch_on(int device) {
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    CC4S_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    CC4S_ASSERT(ptr != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[device] = ptr;
  }
  return g_scratch_ptr[device];
}

}  // namespace

extern "C" {

int cc4s_ccsd_energy_noalloc(int field, int64_t No, int64_t Nv,
                             const void *T_device, const void *Vpphh_device,
                             void *stream, double *out) {
  using namespace cc4s_amd;
  try {
    CC4S_ASSERT(T_device && Vpphh_device && out, "null argument");
    CC4S_ASSERT(field == CC4S_FIELD_REAL || field == CC4S_FIELD_COMPLEX,
                "unsupported tensor type as 'operator'");
    CC4S_ASSERT(No > 0 && Nv > 0, "No and Nv must be positive");
    // the kernels index the amplitude set with int64
    CC4S_ASSERT(Nv <= (int64_t(1) << 20) && No <= (int64_t(1) << 20) &&
                    Nv * Nv <= INT64_MAX / (16 * No * No),
                "No, Nv too large: the amplitude set exceeds int64 byte indexing");
    CC4S_ASSERT(energy_partials_size() <= kScratchBytes,
                "energy reduction needs " + std::to_string(energy_partials_size()) +
                    " bytes of scratch, more than the " + std::to_string(kScratchBytes) +
                    " the library holds per device");
    int device = 0;
    CC4S_HIP(hipGetDevice(&device));
    CC4S_ASSERT(device >= 0 && device < kMaxDevices, "device ordinal out of range");
    void *partials = scratch_on(device);
    const Field f = field == CC4S_FIELD_COMPLEX ? Field::Complex : Field::Real;
    cplx D, X;
    energy(static_cast<hipStream_t>(stream), f, T_device, Vpphh_device, No, Nv,
           partials, &D, &X);
    const cplx E = D + X;
    out[0] = E.real(); out[1] = E.imag();
    out[2] = D.real(); out[3] = D.imag();
    out[4] = X.real(); out[5] = X.imag();
    return 0;
  } catch (const std::exception &e) {
    set_last_error(e.what());
  } catch (...) {
    set_last_error("unknown error");
  }
  return 1;
}

}  // extern "C"
