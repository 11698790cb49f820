// Fused particle-particle ladder (FP64 complex): the V tiles of
// Ccsd.cxx:579-600 are formed from the dressed vertices inside the kernel
// that contracts them with tau, so V never reaches HBM.
#pragma once

#include "tensor.hpp"

namespace cc4s_amd {

// One fused-ladder launch for all slice pairs m <= n of size `slice`
// (Ccsd.cxx:579-600):
//   V_xcyd = sum_G op(left)[G,x,c] op(right)[G,y,d]  (x in slice n, y in slice m)
//   R2[a+x, b+y, i, j] += sum_cd V_xcyd tau[c,d,i,j], and for n > m also
//   R2[b+y, a+x, j, i] += the same.
// `left`, `right` are [G, Nv, Nv] views (any strides, conjugation flags
// honoured); `tau` and `R2` are [Nv, Nv, No, No] (tau dense in (i, j)).
// Returns the algorithmic flops (8 per complex MAC).
struct FusedLadderTable {
  DeviceBuffer blocks;      // pair blocks on the device
  int64_t v = -1, slice = -1;
  int count = 0;
};
double ladder_fused(const View &left, const View &right, const View &tau,
                    const View &R2, int64_t slice, FusedLadderTable &table,
                    hipStream_t s);

}  // namespace cc4s_amd
