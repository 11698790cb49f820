// Paired particle-particle ladder: the slab contraction of Ccsd.cxx:590 over
// pairs c <= d and i <= j (complex).
//
// tau has the pair symmetry tau_cdij = tau_dcji (T2 and T1 x T1 both have
// it), so with
//   tau+_cd(ij) = (tau_cdij + tau_cdji) / 2   symmetric in (c, d),
//   tau-_cd(ij) = (tau_cdij - tau_cdji) / 2   antisymmetric in (c, d),
// the slab contraction R_xyij = sum_cd V_xycd tau_cdij becomes
//   Rp_xy(ij) = sum_{c<=d} Vp_xy(cd) tau+_(cd)(ij),
//               Vp = V_xycd + V_xydc (c < d), V_xycc (c = d);
//   Rm_xy(ij) = sum_{c<d}  Vm_xy(cd) tau-_(cd)(ij),  Vm = V_xycd - V_xydc;
//   R_xyij = Rp + Rm and R_xyji = Rp - Rm (i < j), R_xyii = Rp_xy(ii):
// two contractions over about Nv^2/2 k and No^2/2 columns each, half the
// multiply-adds of the one over Nv^2 k and No^2 columns.
//
// Triangular pair indices (column-major upper triangles): (c, d), c <= d ->
// c + d(d+1)/2; c < d -> c + d(d-1)/2; the same for (i, j).  Layouts, k
// fastest: tau+ [Kp][Np], tau- [Km][Nm], Vp [Kp][x][y], Vm [Km][x][y]; Rp
// [x][y][Np], Rm [x][y][Nm].
#include "ladder.hpp"

#include <algorithm>
#include <cmath>

namespace cc4s_amd {

namespace {

// (i, j) of the upper-triangle pair index n (i <= j)
__device__ __forceinline__ void tri_pair(int64_t n, int64_t &i, int64_t &j) {
  int64_t jj = static_cast<int64_t>((sqrt(8.0 * static_cast<double>(n) + 1.0) - 1.0) * 0.5);
  while (jj * (jj + 1) / 2 > n) --jj;
  while ((jj + 1) * (jj + 2) / 2 <= n) ++jj;
  j = jj;
  i = n - jj * (jj + 1) / 2;
}

__device__ __forceinline__ double2 add2(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 sub2(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}

// grid (ceil(v/256), v, Np): thread c, block row d, block column n = (i, j)
__global__ __launch_bounds__(256) void tau_pairs_kernel(const double2 *tau, double2 *Tp,
                                                        double2 *Tm, int64_t v, int64_t o,
                                                        int64_t Kp, int64_t Km) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  const int64_t d = blockIdx.y;
  if (c > d) return;
  int64_t i, j;
  tri_pair(blockIdx.z, i, j);
  const int64_t vv = v * v;
  const double2 t1 = tau[c + v * d + vv * (i + o * j)];
  const double2 t2 = tau[c + v * d + vv * (j + o * i)];
  const int64_t n = blockIdx.z;
  Tp[c + d * (d + 1) / 2 + Kp * n] = make_double2(0.5 * (t1.x + t2.x), 0.5 * (t1.y + t2.y));
  if (c < d && i < j)
    Tm[c + d * (d - 1) / 2 + Km * (i + j * (j - 1) / 2)] =
        make_double2(0.5 * (t1.x - t2.x), 0.5 * (t1.y - t2.y));
}

constexpr int kT = 32;  // (c, d) tile of the vertex fold

// one 32 x 32 tile (cb <= db) of the (c, d) plane of one (x, y): both
// orientations are read along their contiguous first index and meet in LDS.
// grid (tiles, cx * Ny)
__global__ __launch_bounds__(256) void vertex_pairs_kernel(const double2 *slab, double2 *Vp,
                                                           double2 *Vm, int64_t v, int64_t cx,
                                                           int64_t Kp, int64_t Km) {
  __shared__ double2 t1[kT][kT + 1];  // [d][c]: V_xycd
  __shared__ double2 t2[kT][kT + 1];  // [c][d]: V_xydc
  int64_t cb, db;
  tri_pair(blockIdx.x, cb, db);
  const int64_t xy = blockIdx.y;
  const int64_t x = xy % cx, y = xy / cx;
  const int64_t c0 = cb * kT, d0 = db * kT;
  // slab[c, x, d, y] at c + v (x + cx (d + v y))
  const int64_t sd = v * cx;
  const double2 *base = slab + v * (x + cx * v * y);
  const int t = threadIdx.x;
  for (int e = t; e < kT * kT; e += 256) {
    const int a = e % kT, b = e / kT;
    // V_xycd at c = c0 + a, d = d0 + b; V_xydc at d = d0 + a (first index), c = c0 + b
    const int64_t c1 = c0 + a, d1 = d0 + b;
    t1[b][a] = (c1 < v && d1 < v) ? base[c1 + sd * d1] : make_double2(0.0, 0.0);
    const int64_t d2 = d0 + a, c2 = c0 + b;
    t2[b][a] = (c2 < v && d2 < v) ? base[d2 + sd * c2] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  const int64_t ox = xy;  // (x, y) column of Vp / Vm
  for (int e = t; e < kT * kT; e += 256) {
    const int a = e % kT, b = e / kT;
    const int64_t c = c0 + a, d = d0 + b;
    if (c > d || d >= v) continue;
    const double2 vcd = t1[b][a];
    if (c == d) {
      Vp[c + d * (d + 1) / 2 + Kp * ox] = vcd;
    } else {
      const double2 vdc = t2[a][b];
      Vp[c + d * (d + 1) / 2 + Kp * ox] = add2(vcd, vdc);
      Vm[c + d * (d - 1) / 2 + Km * ox] = sub2(vcd, vdc);
    }
  }
}

// grid (ceil(cx Ny / 256), o * o): R_xyij from Rp and Rm
__global__ __launch_bounds__(256) void unpack_pairs_kernel(const double2 *Rp, const double2 *Rm,
                                                           double2 *R, int64_t nxy, int64_t o) {
  const int64_t xy = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (xy >= nxy) return;
  const int64_t ij = blockIdx.y;
  const int64_t i = ij % o, j = ij / o;
  double2 r;
  if (i == j) {
    r = Rp[xy + nxy * (i + i * (i + 1) / 2)];
  } else if (i < j) {
    r = add2(Rp[xy + nxy * (i + j * (j + 1) / 2)], Rm[xy + nxy * (i + j * (j - 1) / 2)]);
  } else {
    r = sub2(Rp[xy + nxy * (j + i * (i + 1) / 2)], Rm[xy + nxy * (j + i * (i - 1) / 2)]);
  }
  R[xy + nxy * ij] = r;
}

}  // namespace

void ladder_tau_pairs(const void *tau, void *Tp, void *Tm, int64_t v, int64_t o,
                      hipStream_t s) {
  const int64_t Kp = v * (v + 1) / 2, Km = v * (v - 1) / 2, Np = o * (o + 1) / 2;
  CC4S_ASSERT(v > 0 && o > 0 && Np < 65536 && v < 65536, "paired ladder: shape out of range");
  const dim3 grid(static_cast<unsigned>((v + 255) / 256), static_cast<unsigned>(v),
                  static_cast<unsigned>(Np));
  tau_pairs_kernel<<<grid, 256, 0, s>>>(static_cast<const double2 *>(tau),
                                        static_cast<double2 *>(Tp),
                                        static_cast<double2 *>(Tm), v, o, Kp, Km);
  CC4S_HIP(hipGetLastError());
}

void ladder_vertex_pairs(const void *slab, void *Vp, void *Vm, int64_t v, int64_t cx,
                         int64_t Ny, hipStream_t s) {
  const int64_t Kp = v * (v + 1) / 2, Km = v * (v - 1) / 2;
  const int64_t nb = (v + kT - 1) / kT;
  CC4S_ASSERT(cx * Ny < 65536, "paired ladder: slab chunk out of range");
  const dim3 grid(static_cast<unsigned>(nb * (nb + 1) / 2), static_cast<unsigned>(cx * Ny));
  vertex_pairs_kernel<<<grid, 256, 0, s>>>(static_cast<const double2 *>(slab),
                                           static_cast<double2 *>(Vp),
                                           static_cast<double2 *>(Vm), v, cx, Kp, Km);
  CC4S_HIP(hipGetLastError());
}

void ladder_unpack_pairs(const void *Rp, const void *Rm, void *R, int64_t nxy, int64_t o,
                         hipStream_t s) {
  CC4S_ASSERT(o * o < 65536, "paired ladder: No out of range");
  const dim3 grid(static_cast<unsigned>((nxy + 255) / 256), static_cast<unsigned>(o * o));
  unpack_pairs_kernel<<<grid, 256, 0, s>>>(static_cast<const double2 *>(Rp),
                                           static_cast<const double2 *>(Rm),
                                           static_cast<double2 *>(R), nxy, o);
  CC4S_HIP(hipGetLastError());
}

}  // namespace cc4s_amd
