// Energy kernel of the (T) correction.  For one hole pair (i, j) and a chunk
// of k, W[a,b,c,k] holds the six-fold symmetrised connected triples
// amplitude W_abcijk (PerturbativeTriplesReference.cxx:106-116, formed by
// contractions on the MFMA engine); this kernel adds the disconnected part,
// the spin-adapted combination and the denominator (:118-163):
//   S_abc = W_abc + T1_ai V_bcjk + T1_bj V_acik + T1_ck V_abij
//   Z     = 4/3 S_abc - 2 S_acb + 2/3 S_bca
//   E    += Re(W_abc Z) / (e_i + e_j + e_k - e_a - e_b - e_c)
// HBM-bound: W is read at three permutations per element.  Each block keeps
// its running sum across launches in `acc[block]` (one stream: a fixed order),
// so the energy is reproducible run to run.
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
    const T wabc = W[e];
    const T wacb = W[base + a + v * (c + v * b)];
    const T wbca = W[base + b + v * (c + v * a)];
    const T t_ai = T1[a + v * i], t_bj = T1[b + v * j], t_ck = T1[c + v * k];
    const T t_cj = T1[c + v * j], t_bk = T1[b + v * k];
    const T t_bi = T1[b + v * i], t_ak = T1[a + v * k];
    const T s_abc = S::plus(wabc, S::plus(S::mul(t_ai, Vx(b, c, j, k)),
                                          S::plus(S::mul(t_bj, Vx(a, c, i, k)),
                                                  S::mul(t_ck, Vx(a, b, i, j)))));
    const T s_acb = S::plus(wacb, S::plus(S::mul(t_ai, Vx(c, b, j, k)),
                                          S::plus(S::mul(t_cj, Vx(a, b, i, k)),
                                                  S::mul(t_bk, Vx(a, c, i, j)))));
    const T s_bca = S::plus(wbca, S::plus(S::mul(t_bi, Vx(c, a, j, k)),
                                          S::plus(S::mul(t_cj, Vx(b, a, i, k)),
                                                  S::mul(t_ak, Vx(b, c, i, j)))));
    const T z = S::lin(4.0 / 3.0, s_abc, -2.0, s_acb, 2.0 / 3.0, s_bca);
    const double d = ((eh[i] + eh[j]) + eh[k]) - ((ep[a] + ep[b]) + ep[c]);
    sum += S::re_mul(wabc, z) / d;
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
  CC4S_ASSERT(i >= 0 && i < No && j >= 0 && j < No && k0 >= 0 && kc > 0 && k0 + kc <= No,
              "triples: hole indices out of range");
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
