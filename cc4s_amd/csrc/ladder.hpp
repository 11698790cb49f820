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

// Paired ladder (complex; ladder_pairs.hip): with tau_cdij = tau_dcji the
// slab contraction R_xyij = sum_cd V_xycd tau_cdij splits into
//   Rp_xy(ij) = sum_{c<=d} Vp_xy(cd) tau+_(cd)(ij),
//   Rm_xy(ij) = sum_{c<d}  Vm_xy(cd) tau-_(cd)(ij),
// tau+- = (tau_cdij +- tau_cdji)/2, Vp = V_xycd + V_xydc (V_xycc on the
// diagonal), Vm = V_xycd - V_xydc; R_xyij = Rp + Rm, R_xyji = Rp - Rm: half
// the multiply-adds.  Pair indices c <= d -> c + d(d+1)/2, c < d ->
// c + d(d-1)/2 (the same for i, j).
//   ladder_tau_pairs:    tau [c,d,i,j] dense -> tau+ [Kp][Np], tau- [Km][Nm]
//   ladder_vertex_pairs: slab [c,x,d,y] dense (v, cx, v, Ny) -> Vp [Kp][x][y],
//                        Vm [Km][x][y]
//   ladder_unpack_pairs: Rp [xy][Np], Rm [xy][Nm] -> R [xy][i][j]
// with Kp = v(v+1)/2, Km = v(v-1)/2, Np = o(o+1)/2, Nm = o(o-1)/2, k or xy
// fastest.
void ladder_tau_pairs(const void *tau, void *Tp, void *Tm, int64_t v, int64_t o,
                      hipStream_t s);
void ladder_vertex_pairs(const void *slab, void *Vp, void *Vm, int64_t v, int64_t cx,
                         int64_t Ny, hipStream_t s);
void ladder_unpack_pairs(const void *Rp, const void *Rm, void *R, int64_t nxy, int64_t o,
                         hipStream_t s);

}  // namespace cc4s_amd
