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
// 1/6 for three equal.  HBM-bound: W is read at six permutations per
// element.  Each block keeps its running sum across launches in
// `acc[block]` (one stream: a fixed order), so the energy is reproducible.
#include "triples.hpp"

namespace cc4s_amd {

namespace {

constexpr int kBlock = 256;

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

template <bool CPLX>
__global__ __launch_bounds__(kBlock) void triples_energy_kernel(
    const typename TS<CPLX>::T *__restrict__ W, const typename TS<CPLX>::T *__restrict__ V,
    const typename TS<CPLX>::T *__restrict__ T1, const double *__restrict__ eh,
    const double *__restrict__ ep, int64_t No, int64_t Nv, int64_t i, int64_t j,
    int64_t k0, int64_t kc, double *acc) {
  using S = TS<CPLX>;
  using T = typename S::T;
  const int64_t v = Nv, o = No;
  const int64_t n = v * v * v * kc;
  // V_pphh[p,q,s,r] at p + v (q + v (s + o r)); T1[a,i] at a + v i
  auto Vx = [&](int64_t p, int64_t q, int64_t s, int64_t r) {
    return V[p + v * (q + v * (s + o * r))];
  };
  double sum = 0.0;
  for (int64_t e = blockIdx.x * int64_t(kBlock) + threadIdx.x; e < n;
       e += int64_t(gridDim.x) * kBlock) {
    const int64_t a = e % v, b = (e / v) % v, c = (e / (v * v)) % v, kl = e / (v * v * v);
    const int64_t k = k0 + kl;
    const int64_t base = v * v * v * kl;
    // S at the particle tuple (x1, x2, x3) and the holes (i, j, k)
    auto Sx = [&](int64_t x1, int64_t x2, int64_t x3) {
      const T w = W[base + x1 + v * (x2 + v * x3)];
      return S::plus(w, S::plus(S::mul(T1[x1 + v * i], Vx(x2, x3, j, k)),
                                S::plus(S::mul(T1[x2 + v * j], Vx(x1, x3, i, k)),
                                        S::mul(T1[x3 + v * k], Vx(x1, x2, i, j)))));
    };
    const T zs = S::lin(8.0, Sx(a, b, c),
                        -4.0, S::plus(Sx(b, a, c), S::plus(Sx(c, b, a), Sx(a, c, b))),
                        2.0, S::plus(Sx(b, c, a), Sx(c, a, b)));
    const double weight = (i == j && j == k) ? 1.0 / 6.0 : (i == j || j == k) ? 0.5 : 1.0;
    const double d = ((eh[i] + eh[j]) + eh[k]) - ((ep[a] + ep[b]) + ep[c]);
    sum += weight * S::re_mul(W[e], zs) / d;
  }
  __shared__ double sh[kBlock];
  sh[threadIdx.x] = sum;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) acc[blockIdx.x] += sh[0];
}

}  // namespace

void triples_energy(hipStream_t s, Field f, const void *W, const void *Vpphh,
                    const void *T1, const double *eps_h, const double *eps_p,
                    int64_t No, int64_t Nv, int64_t i, int64_t j, int64_t k0,
                    int64_t kc, double *acc) {
  CC4S_ASSERT(i >= 0 && i <= j && j < No && k0 >= j && kc > 0 && k0 + kc <= No,
              "triples: hole indices out of range (needs i <= j <= k)");
  if (f == Field::Complex)
    triples_energy_kernel<true><<<kTriplesBlocks, kBlock, 0, s>>>(
        static_cast<const double2 *>(W), static_cast<const double2 *>(Vpphh),
        static_cast<const double2 *>(T1), eps_h, eps_p, No, Nv, i, j, k0, kc, acc);
  else
    triples_energy_kernel<false><<<kTriplesBlocks, kBlock, 0, s>>>(
        static_cast<const double *>(W), static_cast<const double *>(Vpphh),
        static_cast<const double *>(T1), eps_h, eps_p, No, Nv, i, j, k0, kc, acc);
  CC4S_HIP(hipGetLastError());
}

}  // namespace cc4s_amd
