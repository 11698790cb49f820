// Energy kernel of the (T) correction.  For one hole pair (i, j) and a chunk
// of k, W[a,b,c,k] holds the six-fold symmetrised connected triples
// amplitude W_abcijk (PerturbativeTriplesReference.cxx:106-116, formed by
// contractions on the MFMA engine); this kernel adds the disconnected part,
// the spin-adapted combination and the denominator (:118-163):
//   S_x   = W_x + T1_{x1 i} V_{x2 x3 j k} + T1_{x2 j} V_{x1 x3 i k} + T1_{x3 k} V_{x1 x2 i j}
//   Z     = 4/3 S_abc - 2 S_acb + 2/3 S_bca
//   E    += Re(W_abc Z) / (e_i + e_j + e_k - e_a - e_b - e_c).
// W and S are unchanged by a simultaneous permutation of the (a,i), (b,j),
// (c,k) pairs, so the six orderings of a hole triple contribute together
//   Re(W_abc Zs) / D,  Zs = 8 S_abc - 4 (S_bac + S_cba + S_acb) + 2 (S_bca + S_cab)
// (the sum of Z over the six orderings, written at the ordering i <= j <= k);
// i <= j <= k runs once, weighted 1 for distinct holes, 1/2 for two equal,
// 1/6 for three equal.  Bandwidth-bound: each element of S is formed once,
// in LDS, and read there at its six orderings.  Each block keeps its running
// sum across launches in `acc[block]` (one stream: a fixed order), so the
// energy is reproducible.  Blocks of 512 threads, as many as are resident at
// once (the grid is one wave of blocks: a second, partial wave would leave
// most CUs idle while it runs).
#include "triples.hpp"

namespace cc4s_amd {

namespace {

constexpr int kBlock = 512;
constexpr int kComplexBlocks = 768;

template <bool CPLX>
struct TS;
template <>
struct TS<false> {
  using T = double;
  __device__ static T mul(T a, T b) { return a * b; }
  __device__ static double re_mul(T a, T b) { return a * b; }
  __device__ static T lin(double x, T a, double y, T b, double z, T c) {
    return x * a + y * b + z * c;
  }
  __device__ static T plus(T a, T b) { return a + b; }
};
template <>
struct TS<true> {
  using T = double2;
  __device__ static T mul(T a, T b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
  }
  __device__ static double re_mul(T a, T b) { return a.x * b.x - a.y * b.y; }
  __device__ static T lin(double x, T a, double y, T b, double z, T c) {
    return make_double2(x * a.x + y * b.x + z * c.x, x * a.y + y * b.y + z * c.y);
  }
  __device__ static T plus(T a, T b) { return make_double2(a.x + b.x, a.y + b.y); }
};

// The particle triples are taken in tiles of kTile^3 (a, b, c).  A block
// takes a canonical tile triple tA <= tB <= tC and forms S at the six
// orderings of its three index ranges in LDS, reading W, V and T1 with x1
// fastest (coalesced rows); every distinct permutation of the tile triple is
// then evaluated from those six LDS tiles, so each element of S is formed
// once and every permuted read of it comes from LDS.
constexpr int kTile = 8;
constexpr int kTileElems = kTile * kTile * kTile;

// ordering p = abc, bac, cba, acb, bca, cab: position d of (x1, x2, x3) holds
// particle kOrder[p][d] of (a, b, c); kCoef[p] is its weight in Zs
__constant__ int kOrder[6][3] = {{0, 1, 2}, {1, 0, 2}, {2, 1, 0},
                                 {0, 2, 1}, {1, 2, 0}, {2, 0, 1}};
__constant__ double kCoef[6] = {8.0, -4.0, -4.0, -4.0, 2.0, 2.0};
// kCompose[s][p] = q with kOrder[q][d] = kOrder[s][kOrder[p][d]]: in the
// output tile whose (a, b, c) ranges are the canonical ranges kOrder[s],
// ordering p reads the LDS tile formed for ordering q
__constant__ int kCompose[6][6] = {{0, 1, 2, 3, 4, 5}, {1, 0, 5, 4, 3, 2},
                                   {2, 4, 0, 5, 1, 3}, {3, 5, 4, 0, 2, 1},
                                   {4, 2, 3, 1, 5, 0}, {5, 3, 1, 2, 0, 4}};

template <bool CPLX>
__global__ __launch_bounds__(kBlock) void triples_energy_kernel(
    const typename TS<CPLX>::T *__restrict__ W, const typename TS<CPLX>::T *__restrict__ V,
    const typename TS<CPLX>::T *__restrict__ T1, const double *__restrict__ eh,
    const double *__restrict__ ep, int64_t No, int64_t Nv, int64_t i, int64_t j,
    int64_t k0, int64_t kc, double *acc) {
  using S = TS<CPLX>;
  using T = typename S::T;
  // S at the six orderings of the canonical tile triple's index ranges
  __shared__ T tile[6][kTileElems];
  __shared__ double sh[kBlock / 64];
  const int64_t v = Nv, o = No;
  const int64_t nt = (v + kTile - 1) / kTile;
  const int64_t ntiles = nt * nt * nt * kc;
  // V_pphh[p,q,s,r] at p + v (q + v (s + o r)); T1[a,i] at a + v i
  auto Vx = [&](int64_t p, int64_t q, int64_t s, int64_t r) {
    return V[p + v * (q + v * (s + o * r))];
  };
  T zero;
  if constexpr (CPLX) zero = make_double2(0.0, 0.0); else zero = 0.0;
  double sum = 0.0;
  for (int64_t tid = blockIdx.x; tid < ntiles; tid += gridDim.x) {
    const int64_t ti[3] = {tid % nt, (tid / nt) % nt, (tid / (nt * nt)) % nt};
    if (ti[0] > ti[1] || ti[1] > ti[2]) continue;  // not canonical (uniform per block)
    const int64_t r0[3] = {ti[0] * kTile, ti[1] * kTile, ti[2] * kTile};
    const int64_t kl = tid / (nt * nt * nt);
    const int64_t base = v * v * v * kl;
    const int64_t k = k0 + kl;
    __syncthreads();  // the previous tile's LDS reads are done
    // S at the ordering (x1, x2, x3) and the holes (i, j, k): W plus the
    // disconnected part; x1 runs fastest, so the V and T1 reads along x1 are
    // coalesced and the others are broadcasts
    for (int e = threadIdx.x; e < 6 * kTileElems; e += kBlock) {
      const int p = e / kTileElems, t = e % kTileElems;
      const int64_t x1 = r0[kOrder[p][0]] + t % kTile;
      const int64_t x2 = r0[kOrder[p][1]] + (t / kTile) % kTile;
      const int64_t x3 = r0[kOrder[p][2]] + t / (kTile * kTile);
      T sv = zero;
      if (x1 < v && x2 < v && x3 < v)
        sv = S::plus(W[base + x1 + v * (x2 + v * x3)],
                     S::plus(S::mul(T1[x1 + v * i], Vx(x2, x3, j, k)),
                             S::plus(S::mul(T1[x2 + v * j], Vx(x1, x3, i, k)),
                                     S::mul(T1[x3 + v * k], Vx(x1, x2, i, j)))));
      tile[p][t] = sv;
    }
    __syncthreads();
    const double weight = (i == j && j == k) ? 1.0 / 6.0 : (i == j || j == k) ? 0.5 : 1.0;
    for (int e = threadIdx.x; e < 6 * kTileElems; e += kBlock) {
      const int sg = e / kTileElems, t = e % kTileElems;
      // output tile: (a, b, c) over the canonical ranges kOrder[sg]; a
      // permutation that repeats an earlier tile triple is skipped
      bool repeat = false;
      for (int s2 = 0; s2 < sg; ++s2)
        repeat = repeat || (ti[kOrder[s2][0]] == ti[kOrder[sg][0]] &&
                            ti[kOrder[s2][1]] == ti[kOrder[sg][1]] &&
                            ti[kOrder[s2][2]] == ti[kOrder[sg][2]]);
      if (repeat) continue;
      const int l[3] = {t % kTile, (t / kTile) % kTile, t / (kTile * kTile)};
      const int64_t x[3] = {r0[kOrder[sg][0]] + l[0], r0[kOrder[sg][1]] + l[1],
                            r0[kOrder[sg][2]] + l[2]};
      if (x[0] >= v || x[1] >= v || x[2] >= v) continue;
      T zs = zero;
#pragma unroll
      for (int p = 0; p < 6; ++p) {
        const int q0 = kOrder[p][0], q1 = kOrder[p][1], q2 = kOrder[p][2];
        zs = S::lin(1.0, zs, kCoef[p],
                    tile[kCompose[sg][p]][l[q0] + kTile * (l[q1] + kTile * l[q2])], 0.0, zero);
      }
      const double d = ((eh[i] + eh[j]) + eh[k]) - ((ep[x[0]] + ep[x[1]]) + ep[x[2]]);
      const T w = W[base + x[0] + v * (x[1] + v * x[2])];
      sum += weight * S::re_mul(w, zs) / d;
    }
  }
  // wavefront sums, then the wave sums in wave order
  sum = wave_sum(sum);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = sh[0];
    for (int k = 1; k < kBlock / 64; ++k) b += sh[k];
    acc[blockIdx.x] += b;
  }
}

}  // namespace

void triples_energy(hipStream_t s, Field f, const void *W, const void *Vpphh,
                    const void *T1, const double *eps_h, const double *eps_p,
                    int64_t No, int64_t Nv, int64_t i, int64_t j, int64_t k0,
                    int64_t kc, double *acc) {
  CC4S_ASSERT(i >= 0 && i <= j && j < No && k0 >= j && kc > 0 && k0 + kc <= No,
              "triples: hole indices out of range (needs i <= j <= k)");
  if (f == Field::Complex)
    triples_energy_kernel<true><<<kComplexBlocks, kBlock, 0, s>>>(
        static_cast<const double2 *>(W), static_cast<const double2 *>(Vpphh),
        static_cast<const double2 *>(T1), eps_h, eps_p, No, Nv, i, j, k0, kc, acc);
  else
    triples_energy_kernel<false><<<kTriplesBlocks, kBlock, 0, s>>>(
        static_cast<const double *>(W), static_cast<const double *>(Vpphh),
        static_cast<const double *>(T1), eps_h, eps_p, No, Nv, i, j, k0, kc, acc);
  CC4S_HIP(hipGetLastError());
}

}  // namespace cc4s_amd
