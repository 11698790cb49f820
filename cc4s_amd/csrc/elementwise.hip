// Bandwidth-bound kernels.  Reductions are two-pass with a fixed grid, so
// every result is bitwise reproducible run to run.
#include "elementwise.hpp"

#include <cstring>

namespace cc4s_amd {

namespace {

constexpr int kBlock = 256;
constexpr int kReduceBlocks = 1024;  // fixed: deterministic reduction order
constexpr int kMaxDiis = 16;

inline int grid_for(int64_t n) {
  return static_cast<int>(std::min<int64_t>((n + kBlock - 1) / kBlock, 256 * 32));
}

template <bool CPLX>
struct Sc;
template <>
struct Sc<false> {
  using T = double;
  __device__ static T mul(T a, T b) { return a * b; }
  __device__ static T add(T a, T b) { return a + b; }
  __device__ static T scale(T a, double s) { return a * s; }
  __device__ static T zero() { return 0.0; }
  // conj(a) * b
  __device__ static double2 cdot(T a, T b) { return make_double2(a * b, 0.0); }
};
template <>
struct Sc<true> {
  using T = double2;
  __device__ static T mul(T a, T b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
  }
  __device__ static T add(T a, T b) { return make_double2(a.x + b.x, a.y + b.y); }
  __device__ static T scale(T a, double s) { return make_double2(a.x * s, a.y * s); }
  __device__ static T zero() { return make_double2(0.0, 0.0); }
  __device__ static double2 cdot(T a, T b) {
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
  }
};

template <bool CPLX>
__global__ __launch_bounds__(kBlock) void tau_kernel(typename Sc<CPLX>::T *X,
                                                     const typename Sc<CPLX>::T *T2,
                                                     const typename Sc<CPLX>::T *T1,
                                                     double f, int64_t No, int64_t Nv) {
  using S = Sc<CPLX>;
  const int64_t n = Nv * Nv * No * No;
  for (int64_t e = blockIdx.x * int64_t(kBlock) + threadIdx.x; e < n;
       e += int64_t(gridDim.x) * kBlock) {
    int64_t r = e;
    const int64_t a = r % Nv; r /= Nv;
    const int64_t b = r % Nv; r /= Nv;
    const int64_t i = r % No;
    const int64_t j = r / No;
    X[e] = S::add(T2[e], S::scale(S::mul(T1[a + Nv * i], T1[b + Nv * j]), f));
  }
}

// (a, b) tiles of the pphh layout: a tile at fixed (i, j) and its partner
// (the transposed tile at (j, i)) move through LDS together, so both are
// read and written along their contiguous index
constexpr int kPairTile = 32;

// R_abij <- R_abij + R_baji (old values on the right): one block per
// unordered pair of tiles {(A, B, i, j), (B, A, j, i)}, taken where
// (i, tile A) <= (j, tile B); both tiles are read before either is written
template <bool CPLX>
__global__ __launch_bounds__(kBlock) void symmetrize_kernel(typename Sc<CPLX>::T *R,
                                                            int64_t No, int64_t Nv) {
  using S = Sc<CPLX>;
  using T = typename S::T;
  constexpr int kt = kPairTile;
  __shared__ T xs[kt][kt + 1], ys[kt][kt + 1];
  const int64_t nt = (Nv + kt - 1) / kt;
  const int64_t ntiles = nt * nt * No * No;
  for (int64_t tid = blockIdx.x; tid < ntiles; tid += gridDim.x) {
    const int64_t ta = tid % nt, tb = (tid / nt) % nt, ij = tid / (nt * nt);
    const int64_t i = ij % No, j = ij / No;
    const int64_t P = i * nt + ta, Q = j * nt + tb;
    if (P > Q) continue;  // the partner's block does it (uniform per block)
    const int64_t a0 = ta * kt, b0 = tb * kt;
    const int64_t offX = Nv * Nv * (i + No * j), offY = Nv * Nv * (j + No * i);
    __syncthreads();  // the previous pair's LDS reads are done
    for (int e = threadIdx.x; e < kt * kt; e += kBlock) {
      const int u = e % kt, w = e / kt;
      // xs[a - a0][b - b0] = R_abij, ys[b - b0][a - a0] = R_baji
      xs[u][w] = (a0 + u < Nv && b0 + w < Nv) ? R[offX + (a0 + u) + Nv * (b0 + w)] : S::zero();
      ys[u][w] = (b0 + u < Nv && a0 + w < Nv) ? R[offY + (b0 + u) + Nv * (a0 + w)] : S::zero();
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kt * kt; e += kBlock) {
      const int u = e % kt, w = e / kt;
      if (a0 + u < Nv && b0 + w < Nv)
        R[offX + (a0 + u) + Nv * (b0 + w)] = S::add(xs[u][w], ys[w][u]);
      if (P != Q && b0 + u < Nv && a0 + w < Nv)
        R[offY + (b0 + u) + Nv * (a0 + w)] = S::add(ys[u][w], xs[w][u]);
    }
  }
}

template <bool CPLX>
__global__ __launch_bounds__(kBlock) void update_kernel(
    const typename Sc<CPLX>::T *R, const typename Sc<CPLX>::T *T,
    const double *eh, const double *ep, double shift, typename Sc<CPLX>::T *Tn,
    typename Sc<CPLX>::T *dT, int64_t No, int64_t Nv) {
  using S = Sc<CPLX>;
  const int64_t n1 = Nv * No;
  const int64_t n = n1 + Nv * Nv * No * No;
  for (int64_t e = blockIdx.x * int64_t(kBlock) + threadIdx.x; e < n;
       e += int64_t(gridDim.x) * kBlock) {
    double delta;
    if (e < n1) {
      const int64_t a = e % Nv, i = e / Nv;
      delta = (0.0 + ep[a]) - eh[i];
    } else {
      int64_t r = e - n1;
      const int64_t a = r % Nv; r /= Nv;
      const int64_t b = r % Nv; r /= Nv;
      const int64_t i = r % No;
      const int64_t j = r / No;
      delta = (((0.0 + ep[a]) - eh[i]) + ep[b]) - eh[j];
    }
    typename S::T rv = R[e];
    if (T && shift != 0.0) rv = S::add(rv, S::scale(T[e], -shift));
    const double d = -1.0 / (delta + shift);
    const typename S::T tn = S::scale(rv, d);
    Tn[e] = tn;
    dT[e] = T ? S::add(tn, S::scale(T[e], -1.0)) : tn;
  }
}

// block-level sum of double2 values held by every thread: a wavefront
// reduction per wave, then the kBlock/64 wave sums in wave order; every
// thread gets the total
__device__ double2 block_sum(double2 v, double2 *sh) {
  v.x = wave_sum(v.x);
  v.y = wave_sum(v.y);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double2 r = sh[0];
#pragma unroll
  for (int k = 1; k < kBlock / 64; ++k) {
    r.x += sh[k].x;
    r.y += sh[k].y;
  }
  __syncthreads();  // sh may be reused by the next call
  return r;
}

// energy partials: per block (D, X) as two double2.  One work item is an
// unordered pair of (a, b) tiles {(A, B), (B, A)} at fixed (i, j): both V
// tiles move through LDS, so every V_abij and T2_abij is read once, along
// its contiguous index, and serves the direct term of its own element and
// the exchange term of the mirrored one (the tile A = B is a pair by itself)
template <bool CPLX>
__global__ __launch_bounds__(kBlock) void energy_kernel(const typename Sc<CPLX>::T *T,
                                                        const typename Sc<CPLX>::T *V,
                                                        int64_t No, int64_t Nv,
                                                        double2 *partials) {
  using S = Sc<CPLX>;
  using E = typename S::T;
  constexpr int kt = kPairTile;
  __shared__ double2 sh[kBlock];
  // vx[u][w] = V at (a0 + u, b0 + w), vy[u][w] = V at (b0 + u, a0 + w)
  __shared__ E vx[kt][kt + 1], vy[kt][kt + 1];
  const E *T1 = T;
  const E *T2 = T + Nv * No;
  const int64_t nt = (Nv + kt - 1) / kt;
  const int64_t npairs = nt * (nt + 1) / 2;
  const int64_t ntiles = npairs * No * No;
  double2 d = make_double2(0.0, 0.0), x = make_double2(0.0, 0.0);
  for (int64_t tid = blockIdx.x; tid < ntiles; tid += gridDim.x) {
    const int64_t pr = tid % npairs, ij = tid / npairs;
    // pr = tb (tb + 1) / 2 + ta, ta <= tb
    int64_t tb = static_cast<int64_t>((sqrt(8.0 * static_cast<double>(pr) + 1.0) - 1.0) * 0.5);
    while (tb * (tb + 1) / 2 > pr) --tb;
    while ((tb + 1) * (tb + 2) / 2 <= pr) ++tb;
    const int64_t ta = pr - tb * (tb + 1) / 2;
    const bool diag = ta == tb;
    const int64_t i = ij % No, j = ij / No;
    const int64_t a0 = ta * kt, b0 = tb * kt, off = Nv * Nv * (i + No * j);
    __syncthreads();
    for (int e = threadIdx.x; e < kt * kt; e += kBlock) {
      const int u = e % kt, w = e / kt;
      vx[u][w] = (a0 + u < Nv && b0 + w < Nv) ? V[off + (a0 + u) + Nv * (b0 + w)] : S::zero();
      if (!diag)
        vy[u][w] = (b0 + u < Nv && a0 + w < Nv) ? V[off + (b0 + u) + Nv * (a0 + w)] : S::zero();
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kt * kt; e += kBlock) {
      const int u = e % kt, w = e / kt;
      // tau * conj(V) = conj(conj(tau) * V)
      const int64_t a = a0 + u, b = b0 + w;
      if (a < Nv && b < Nv) {
        const E tau = S::add(T2[off + a + Nv * b], S::mul(T1[a + Nv * i], T1[b + Nv * j]));
        const double2 dv = S::cdot(vx[u][w], tau);              // V_abij
        const double2 xv = S::cdot(diag ? vx[w][u] : vy[w][u], tau);  // V_baij
        d.x += dv.x; d.y += dv.y;
        x.x += xv.x; x.y += xv.y;
      }
      const int64_t a2 = b0 + u, b2 = a0 + w;   // element of the tile (B, A)
      if (!diag && a2 < Nv && b2 < Nv) {
        const E tau = S::add(T2[off + a2 + Nv * b2], S::mul(T1[a2 + Nv * i], T1[b2 + Nv * j]));
        const double2 dv = S::cdot(vy[u][w], tau);
        const double2 xv = S::cdot(vx[w][u], tau);
        d.x += dv.x; d.y += dv.y;
        x.x += xv.x; x.y += xv.y;
      }
    }
  }
  const double2 ds = block_sum(d, sh);
  const double2 xs = block_sum(x, sh);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = ds;
    partials[2 * blockIdx.x + 1] = xs;
  }
}

// sums `rows` interleaved series of `nblocks` double2 partials, in order
__global__ __launch_bounds__(kBlock) void finish_kernel(const double2 *partials,
                                                        int rows, int nblocks,
                                                        double2 *out) {
  __shared__ double2 sh[kBlock];
  for (int r = 0; r < rows; ++r) {
    double2 v = make_double2(0.0, 0.0);
    for (int b = threadIdx.x; b < nblocks; b += kBlock) {
      const double2 p = partials[b * rows + r];
      v.x += p.x;
      v.y += p.y;
    }
    const double2 s = block_sum(v, sh);
    if (threadIdx.x == 0) out[r] = s;
  }
}

struct DotArgs {
  const void *x[kMaxDiis];
  const void *A[kMaxDiis];
  const void *R[kMaxDiis];
  double c[kMaxDiis];
};

template <bool CPLX>
__global__ __launch_bounds__(kBlock) void dots_kernel(DotArgs p, int count,
                                                      const typename Sc<CPLX>::T *y,
                                                      int64_t n, double2 *partials) {
  using S = Sc<CPLX>;
  __shared__ double2 sh[kBlock];
  double2 acc[kMaxDiis];
#pragma unroll
  for (int k = 0; k < kMaxDiis; ++k) acc[k] = make_double2(0.0, 0.0);
  for (int64_t e = blockIdx.x * int64_t(kBlock) + threadIdx.x; e < n;
       e += int64_t(gridDim.x) * kBlock) {
    const typename S::T yv = y[e];
#pragma unroll
    for (int k = 0; k < kMaxDiis; ++k) {
      if (k < count) {
        const double2 v = S::cdot(static_cast<const typename S::T *>(p.x[k])[e], yv);
        acc[k].x += v.x;
        acc[k].y += v.y;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kMaxDiis; ++k) {
    if (k < count) {
      const double2 s = block_sum(acc[k], sh);
      if (threadIdx.x == 0) partials[blockIdx.x * count + k] = s;
    }
  }
}

template <bool CPLX>
__global__ __launch_bounds__(kBlock) void combine_kernel(DotArgs p, int count,
                                                         typename Sc<CPLX>::T *out,
                                                         int64_t n, double2 *partials) {
  using S = Sc<CPLX>;
  __shared__ double2 sh[kBlock];
  double2 nrm = make_double2(0.0, 0.0);
  for (int64_t e = blockIdx.x * int64_t(kBlock) + threadIdx.x; e < n;
       e += int64_t(gridDim.x) * kBlock) {
    typename S::T a = S::zero(), r = S::zero();
    for (int k = 0; k < count; ++k) {
      a = S::add(a, S::scale(static_cast<const typename S::T *>(p.A[k])[e], p.c[k]));
      r = S::add(r, S::scale(static_cast<const typename S::T *>(p.R[k])[e], p.c[k]));
    }
    out[e] = a;
    const double2 v = S::cdot(r, r);
    nrm.x += v.x;
  }
  const double2 s = block_sum(nrm, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

}  // namespace

void tau(hipStream_t s, Field f, void *X, const void *T2, const void *T1,
         double factor, int64_t No, int64_t Nv) {
  const int64_t n = Nv * Nv * No * No;
  if (f == Field::Complex)
    tau_kernel<true><<<grid_for(n), kBlock, 0, s>>>(
        static_cast<double2 *>(X), static_cast<const double2 *>(T2),
        static_cast<const double2 *>(T1), factor, No, Nv);
  else
    tau_kernel<false><<<grid_for(n), kBlock, 0, s>>>(
        static_cast<double *>(X), static_cast<const double *>(T2),
        static_cast<const double *>(T1), factor, No, Nv);
  CC4S_HIP(hipGetLastError());
}

void symmetrize_pairs(hipStream_t s, Field f, void *R2, int64_t No, int64_t Nv) {
  if (f == Field::Complex)
    symmetrize_kernel<true><<<kReduceBlocks, kBlock, 0, s>>>(static_cast<double2 *>(R2), No, Nv);
  else
    symmetrize_kernel<false><<<kReduceBlocks, kBlock, 0, s>>>(static_cast<double *>(R2), No, Nv);
  CC4S_HIP(hipGetLastError());
}

void amplitude_update(hipStream_t s, Field f, const void *R, const void *T,
                      const double *eps_h, const double *eps_p, double shift,
                      void *Tnew, void *dT, int64_t No, int64_t Nv) {
  const int64_t n = Nv * No + Nv * Nv * No * No;
  if (f == Field::Complex)
    update_kernel<true><<<grid_for(n), kBlock, 0, s>>>(
        static_cast<const double2 *>(R), static_cast<const double2 *>(T), eps_h,
        eps_p, shift, static_cast<double2 *>(Tnew), static_cast<double2 *>(dT), No, Nv);
  else
    update_kernel<false><<<grid_for(n), kBlock, 0, s>>>(
        static_cast<const double *>(R), static_cast<const double *>(T), eps_h,
        eps_p, shift, static_cast<double *>(Tnew), static_cast<double *>(dT), No, Nv);
  CC4S_HIP(hipGetLastError());
}

size_t energy_partials_size() { return (2 * kReduceBlocks + 2) * sizeof(double2); }

void energy(hipStream_t s, Field f, const void *T, const void *V, int64_t No,
            int64_t Nv, void *partials, cplx *D, cplx *X) {
  double2 *part = static_cast<double2 *>(partials);
  double2 *out = part + 2 * kReduceBlocks;
  if (f == Field::Complex)
    energy_kernel<true><<<kReduceBlocks, kBlock, 0, s>>>(
        static_cast<const double2 *>(T), static_cast<const double2 *>(V), No, Nv, part);
  else
    energy_kernel<false><<<kReduceBlocks, kBlock, 0, s>>>(
        static_cast<const double *>(T), static_cast<const double *>(V), No, Nv, part);
  CC4S_HIP(hipGetLastError());
  finish_kernel<<<1, kBlock, 0, s>>>(part, 2, kReduceBlocks, out);
  CC4S_HIP(hipGetLastError());
  double2 h[2];
  CC4S_HIP(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, s));
  CC4S_HIP(hipStreamSynchronize(s));
  // partial sums hold sum conj(V) tau; the degeneracy factors of
  // CoupledCluster.cxx:224-231 are 0.5*2*2 = 2 (direct) and -0.5*2 = -1
  *D = 2.0 * cplx(h[0].x, h[0].y);
  *X = -1.0 * cplx(h[1].x, h[1].y);
}

size_t dots_partials_size(int count) {
  return (static_cast<size_t>(kReduceBlocks) * count + count) * sizeof(double2);
}

void multi_dot(hipStream_t s, Field f, const void *const *x, int count,
               const void *y, int64_t n, void *partials, std::vector<cplx> &out) {
  CC4S_ASSERT(count >= 1 && count <= kMaxDiis, "too many DIIS vectors");
  DotArgs a;
  for (int k = 0; k < count; ++k) a.x[k] = x[k];
  double2 *part = static_cast<double2 *>(partials);
  double2 *res = part + static_cast<size_t>(kReduceBlocks) * count;
  if (f == Field::Complex)
    dots_kernel<true><<<kReduceBlocks, kBlock, 0, s>>>(
        a, count, static_cast<const double2 *>(y), n, part);
  else
    dots_kernel<false><<<kReduceBlocks, kBlock, 0, s>>>(
        a, count, static_cast<const double *>(y), n, part);
  CC4S_HIP(hipGetLastError());
  finish_kernel<<<1, kBlock, 0, s>>>(part, count, kReduceBlocks, res);
  CC4S_HIP(hipGetLastError());
  std::vector<double2> h(count);
  CC4S_HIP(hipMemcpyAsync(h.data(), res, count * sizeof(double2),
                          hipMemcpyDeviceToHost, s));
  CC4S_HIP(hipStreamSynchronize(s));
  out.resize(count);
  for (int k = 0; k < count; ++k) out[k] = cplx(h[k].x, h[k].y);
}

double combine(hipStream_t s, Field f, const void *const *A, const void *const *R,
               const std::vector<double> &c, int64_t n, void *out, void *partials) {
  const int count = static_cast<int>(c.size());
  CC4S_ASSERT(count >= 1 && count <= kMaxDiis, "too many DIIS vectors");
  DotArgs a;
  for (int k = 0; k < count; ++k) {
    a.A[k] = A[k];
    a.R[k] = R[k];
    a.c[k] = c[k];
  }
  double2 *part = static_cast<double2 *>(partials);
  double2 *res = part + kReduceBlocks;
  if (f == Field::Complex)
    combine_kernel<true><<<kReduceBlocks, kBlock, 0, s>>>(a, count,
                                                          static_cast<double2 *>(out), n, part);
  else
    combine_kernel<false><<<kReduceBlocks, kBlock, 0, s>>>(a, count,
                                                           static_cast<double *>(out), n, part);
  CC4S_HIP(hipGetLastError());
  finish_kernel<<<1, kBlock, 0, s>>>(part, 1, kReduceBlocks, res);
  CC4S_HIP(hipGetLastError());
  double2 h;
  CC4S_HIP(hipMemcpyAsync(&h, res, sizeof(h), hipMemcpyDeviceToHost, s));
  CC4S_HIP(hipStreamSynchronize(s));
  return h.x;
}

// ---------------------------------------------------------------------------
// pair-packed sets
// ---------------------------------------------------------------------------

namespace {

// (P, Q) with P <= Q of pair index pi = Q(Q+1)/2 + P
__device__ __forceinline__ void pair_of(int64_t pi, int64_t &P, int64_t &Q) {
  int64_t q = static_cast<int64_t>((sqrt(8.0 * static_cast<double>(pi) + 1.0) - 1.0) * 0.5);
  while (q * (q + 1) / 2 > pi) --q;
  while ((q + 1) * (q + 2) / 2 <= pi) ++q;
  Q = q;
  P = pi - q * (q + 1) / 2;
}

__device__ __forceinline__ int64_t pair_index(int64_t P, int64_t Q) {
  return P <= Q ? Q * (Q + 1) / 2 + P : P * (P + 1) / 2 + Q;
}

template <bool CPLX>
__global__ __launch_bounds__(kBlock) void update_packed_kernel(
    const typename Sc<CPLX>::T *R, const typename Sc<CPLX>::T *T,
    const double *eh, const double *ep, double shift, typename Sc<CPLX>::T *Tn,
    typename Sc<CPLX>::T *dT, int64_t No, int64_t Nv) {
  using S = Sc<CPLX>;
  const int64_t n1 = Nv * No;
  const int64_t np = n1 * (n1 + 1) / 2;
  for (int64_t t = blockIdx.x * int64_t(kBlock) + threadIdx.x; t < n1 + np;
       t += int64_t(gridDim.x) * kBlock) {
    double delta;
    int64_t e;
    if (t < n1) {
      const int64_t a = t % Nv, i = t / Nv;
      delta = (0.0 + ep[a]) - eh[i];
      e = t;
    } else {
      int64_t P, Q;
      pair_of(t - n1, P, Q);
      const int64_t a = P % Nv, i = P / Nv, b = Q % Nv, j = Q / Nv;
      delta = (((0.0 + ep[a]) - eh[i]) + ep[b]) - eh[j];
      e = n1 + a + Nv * (b + Nv * (i + No * j));
    }
    typename S::T rv = R[e];
    if (T && shift != 0.0) rv = S::add(rv, S::scale(T[e], -shift));
    const typename S::T tn = S::scale(rv, -1.0 / (delta + shift));
    if (Tn) Tn[t] = tn;
    if (dT) dT[t] = T ? S::add(tn, S::scale(T[e], -1.0)) : tn;
  }
}

// packed elements [t0, t0 + len) of one set; x[k] and y point at element t0
template <bool CPLX>
__global__ __launch_bounds__(kBlock) void dots_packed_kernel(DotArgs p, int count,
                                                             const typename Sc<CPLX>::T *y,
                                                             int64_t n1, int64_t t0,
                                                             int64_t len, double2 *partials) {
  using S = Sc<CPLX>;
  __shared__ double2 sh[kBlock];
  double2 acc[kMaxDiis];
#pragma unroll
  for (int k = 0; k < kMaxDiis; ++k) acc[k] = make_double2(0.0, 0.0);
  for (int64_t u = blockIdx.x * int64_t(kBlock) + threadIdx.x; u < len;
       u += int64_t(gridDim.x) * kBlock) {
    const int64_t t = t0 + u;
    double w = 1.0;
    if (t >= n1) {
      int64_t P, Q;
      pair_of(t - n1, P, Q);
      w = P == Q ? 1.0 : 2.0;  // an off-diagonal pair stands for two elements
    }
    const typename S::T yv = y[u];
#pragma unroll
    for (int k = 0; k < kMaxDiis; ++k) {
      if (k < count) {
        const double2 v = S::cdot(static_cast<const typename S::T *>(p.x[k])[u], yv);
        acc[k].x += w * v.x;
        acc[k].y += w * v.y;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kMaxDiis; ++k) {
    if (k < count) {
      const double2 s = block_sum(acc[k], sh);
      if (threadIdx.x == 0) partials[blockIdx.x * count + k] = s;
    }
  }
}

// full-layout output; the norm runs over full elements, i.e. it is the norm
// of the full symmetric tensor.  T2 is taken in (a, b) tiles at fixed
// (i, j); with P = a + Nv i and Q = b + Nv j, a tile below the diagonal of
// the (P, Q) tile grid has its packed index pair_index(P, Q) running along
// b: each pair there is combined once and written to its element in the
// tile (along a, through LDS) and to its mirror in the tile above (along
// b), so every packed element is read once; only the diagonal tiles are
// combined element by element.
template <bool CPLX>
__global__ __launch_bounds__(kBlock) void combine_packed_kernel(DotArgs p, int count,
                                                                typename Sc<CPLX>::T *out,
                                                                int64_t No, int64_t Nv,
                                                                double2 *partials) {
  using S = Sc<CPLX>;
  using T = typename S::T;
  constexpr int kt = kPairTile;
  __shared__ double2 sh[kBlock];
  __shared__ T at[kt][kt + 1];
  const int64_t n1 = Nv * No;
  double2 nrm = make_double2(0.0, 0.0);
  // the combination at packed index t, accumulating |R|^2 of it `mult`
  // times (the number of full-layout elements it is written to)
  auto combine_at = [&](int64_t t, double mult) {
    T av = S::zero(), rv = S::zero();
    for (int k = 0; k < count; ++k) {
      av = S::add(av, S::scale(static_cast<const T *>(p.A[k])[t], p.c[k]));
      rv = S::add(rv, S::scale(static_cast<const T *>(p.R[k])[t], p.c[k]));
    }
    nrm.x += mult * S::cdot(rv, rv).x;
    return av;
  };
  for (int64_t e = blockIdx.x * int64_t(kBlock) + threadIdx.x; e < n1;
       e += int64_t(gridDim.x) * kBlock)
    out[e] = combine_at(e, 1.0);
  const int64_t nt = (Nv + kt - 1) / kt;
  const int64_t ntiles = nt * nt * No * No;
  for (int64_t tid = blockIdx.x; tid < ntiles; tid += gridDim.x) {
    const int64_t ta = tid % nt, tb = (tid / nt) % nt, ij = tid / (nt * nt);
    const int64_t i = ij % No, j = ij / No;
    const int64_t tp = i * nt + ta, tq = j * nt + tb;  // tile ids along P and Q
    if (tp < tq) continue;  // the mirror of a tile below the diagonal (uniform per block)
    const int64_t a0 = ta * kt, b0 = tb * kt;
    const int64_t offX = n1 + Nv * Nv * (i + No * j), offY = n1 + Nv * Nv * (j + No * i);
    if (tp == tq) {
      // diagonal tile (i == j, ta == tb): element by element, both halves
      for (int e = threadIdx.x; e < kt * kt; e += kBlock) {
        const int u = e % kt, w = e / kt;
        const int64_t a = a0 + u, b = b0 + w;
        if (a < Nv && b < Nv)
          out[offX + a + Nv * b] = combine_at(n1 + pair_index(a + Nv * i, b + Nv * j), 1.0);
      }
      continue;
    }
    // every P > Q in the tile: pair_index(P, Q) = P (P + 1) / 2 + Q runs along
    // b.  Each pair is combined once and written to both of its elements:
    // (b, a, j, i) directly along b, (a, b, i, j) along a through LDS.
    __syncthreads();  // the previous tile's LDS reads are done
    for (int e = threadIdx.x; e < kt * kt; e += kBlock) {
      const int u = e % kt, w = e / kt;
      const int64_t a = a0 + w, b = b0 + u;
      if (a < Nv && b < Nv) {
        const T v = combine_at(n1 + pair_index(a + Nv * i, b + Nv * j), 2.0);
        at[u][w] = v;
        out[offY + b + Nv * a] = v;
      }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kt * kt; e += kBlock) {
      const int u = e % kt, w = e / kt;
      const int64_t a = a0 + u, b = b0 + w;
      if (a < Nv && b < Nv) out[offX + a + Nv * b] = at[w][u];
    }
  }
  const double2 s = block_sum(nrm, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

template <bool CPLX>
__global__ __launch_bounds__(kBlock) void unpack_accumulate_kernel(
    const typename Sc<CPLX>::T *x, int64_t t0, int64_t len, double c, int first,
    typename Sc<CPLX>::T *out, int64_t No, int64_t Nv) {
  using S = Sc<CPLX>;
  const int64_t n1 = Nv * No;
  for (int64_t u = blockIdx.x * int64_t(kBlock) + threadIdx.x; u < len;
       u += int64_t(gridDim.x) * kBlock) {
    const int64_t t = t0 + u;
    const typename S::T v = S::scale(x[u], c);
    if (t < n1) {
      out[t] = first ? v : S::add(out[t], v);
      continue;
    }
    int64_t P, Q;
    pair_of(t - n1, P, Q);
    const int64_t a = P % Nv, i = P / Nv, b = Q % Nv, j = Q / Nv;
    const int64_t e = n1 + a + Nv * (b + Nv * (i + No * j));
    out[e] = first ? v : S::add(out[e], v);
    if (P != Q) {
      const int64_t m = n1 + b + Nv * (a + Nv * (j + No * i));
      out[m] = first ? v : S::add(out[m], v);
    }
  }
}

// packed elements [t0, t0 + len); p.R[k] point at element t0
template <bool CPLX>
__global__ __launch_bounds__(kBlock) void norm_packed_kernel(DotArgs p, int count,
                                                             int64_t n1, int64_t t0,
                                                             int64_t len,
                                                             double2 *partials) {
  using S = Sc<CPLX>;
  __shared__ double2 sh[kBlock];
  double nrm = 0.0;
  for (int64_t u = blockIdx.x * int64_t(kBlock) + threadIdx.x; u < len;
       u += int64_t(gridDim.x) * kBlock) {
    const int64_t t = t0 + u;
    double w = 1.0;
    if (t >= n1) {
      int64_t P, Q;
      pair_of(t - n1, P, Q);
      w = P == Q ? 1.0 : 2.0;
    }
    typename S::T rv = S::zero();
    for (int k = 0; k < count; ++k)
      rv = S::add(rv, S::scale(static_cast<const typename S::T *>(p.R[k])[u], p.c[k]));
    nrm += w * S::cdot(rv, rv).x;
  }
  const double2 s = block_sum(make_double2(nrm, 0.0), sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

}  // namespace

namespace {
// y[t] = alpha * x[e(t)] + beta * y[t] for the packed elements t of a set,
// e(t) its full-layout position (the element update_packed_kernel reads)
template <bool CPLX>
__global__ __launch_bounds__(kBlock) void pack_axpby_kernel(
    const typename Sc<CPLX>::T *x, double alpha, typename Sc<CPLX>::T *y, double beta,
    int64_t No, int64_t Nv) {
  using S = Sc<CPLX>;
  const int64_t n1 = Nv * No;
  const int64_t np = n1 * (n1 + 1) / 2;
  for (int64_t t = blockIdx.x * int64_t(kBlock) + threadIdx.x; t < n1 + np;
       t += int64_t(gridDim.x) * kBlock) {
    int64_t e = t;
    if (t >= n1) {
      int64_t P, Q;
      pair_of(t - n1, P, Q);
      const int64_t a = P % Nv, i = P / Nv, b = Q % Nv, j = Q / Nv;
      e = n1 + a + Nv * (b + Nv * (i + No * j));
    }
    const typename S::T v = S::scale(x[e], alpha);
    y[t] = beta == 0.0 ? v : S::add(v, S::scale(y[t], beta));
  }
}
}  // namespace

void pack_axpby(hipStream_t s, Field f, const void *full, double alpha, void *packed,
                double beta, int64_t No, int64_t Nv) {
  const int64_t n = packed_set_size(No, Nv);
  if (f == Field::Complex)
    pack_axpby_kernel<true><<<grid_for(n), kBlock, 0, s>>>(
        static_cast<const double2 *>(full), alpha, static_cast<double2 *>(packed), beta,
        No, Nv);
  else
    pack_axpby_kernel<false><<<grid_for(n), kBlock, 0, s>>>(
        static_cast<const double *>(full), alpha, static_cast<double *>(packed), beta,
        No, Nv);
  CC4S_HIP(hipGetLastError());
}

void unpack_accumulate(hipStream_t s, Field f, const void *chunk, int64_t t0,
                       int64_t len, double c, bool first, void *out, int64_t No,
                       int64_t Nv) {
  if (f == Field::Complex)
    unpack_accumulate_kernel<true><<<grid_for(len), kBlock, 0, s>>>(
        static_cast<const double2 *>(chunk), t0, len, c, first ? 1 : 0,
        static_cast<double2 *>(out), No, Nv);
  else
    unpack_accumulate_kernel<false><<<grid_for(len), kBlock, 0, s>>>(
        static_cast<const double *>(chunk), t0, len, c, first ? 1 : 0,
        static_cast<double *>(out), No, Nv);
  CC4S_HIP(hipGetLastError());
}

double norm_packed(hipStream_t s, Field f, const void *const *R,
                   const std::vector<double> &c, int64_t No, int64_t Nv,
                   void *partials) {
  return norm_packed_range(s, f, R, c, 0, packed_set_size(No, Nv), No, Nv, partials);
}

double norm_packed_range(hipStream_t s, Field f, const void *const *R,
                         const std::vector<double> &c, int64_t t0, int64_t len,
                         int64_t No, int64_t Nv, void *partials) {
  const int count = static_cast<int>(c.size());
  CC4S_ASSERT(count >= 1 && count <= kMaxDiis, "too many DIIS vectors");
  CC4S_ASSERT(t0 >= 0 && len >= 0 && t0 + len <= packed_set_size(No, Nv),
              "packed range out of bounds");
  DotArgs a;
  for (int k = 0; k < count; ++k) {
    a.R[k] = R[k];
    a.c[k] = c[k];
  }
  double2 *part = static_cast<double2 *>(partials);
  double2 *res = part + kReduceBlocks;
  if (f == Field::Complex)
    norm_packed_kernel<true><<<kReduceBlocks, kBlock, 0, s>>>(a, count, Nv * No, t0,
                                                              len, part);
  else
    norm_packed_kernel<false><<<kReduceBlocks, kBlock, 0, s>>>(a, count, Nv * No, t0,
                                                               len, part);
  CC4S_HIP(hipGetLastError());
  finish_kernel<<<1, kBlock, 0, s>>>(part, 1, kReduceBlocks, res);
  CC4S_HIP(hipGetLastError());
  double2 h;
  CC4S_HIP(hipMemcpyAsync(&h, res, sizeof(h), hipMemcpyDeviceToHost, s));
  CC4S_HIP(hipStreamSynchronize(s));
  return h.x;
}

int64_t packed_set_size(int64_t No, int64_t Nv) {
  const int64_t n1 = Nv * No;
  return n1 + n1 * (n1 + 1) / 2;
}

void amplitude_update_packed(hipStream_t s, Field f, const void *R, const void *T,
                             const double *eps_h, const double *eps_p, double shift,
                             void *Tnew, void *dT, int64_t No, int64_t Nv) {
  const int64_t n = packed_set_size(No, Nv);
  if (f == Field::Complex)
    update_packed_kernel<true><<<grid_for(n), kBlock, 0, s>>>(
        static_cast<const double2 *>(R), static_cast<const double2 *>(T), eps_h,
        eps_p, shift, static_cast<double2 *>(Tnew), static_cast<double2 *>(dT), No, Nv);
  else
    update_packed_kernel<false><<<grid_for(n), kBlock, 0, s>>>(
        static_cast<const double *>(R), static_cast<const double *>(T), eps_h,
        eps_p, shift, static_cast<double *>(Tnew), static_cast<double *>(dT), No, Nv);
  CC4S_HIP(hipGetLastError());
}

void multi_dot_packed(hipStream_t s, Field f, const void *const *x, int count,
                      const void *y, int64_t No, int64_t Nv, void *partials,
                      std::vector<cplx> &out) {
  multi_dot_packed_range(s, f, x, count, y, 0, packed_set_size(No, Nv), No, Nv,
                         partials, out);
}

void multi_dot_packed_range(hipStream_t s, Field f, const void *const *x, int count,
                            const void *y, int64_t t0, int64_t len, int64_t No,
                            int64_t Nv, void *partials, std::vector<cplx> &out) {
  CC4S_ASSERT(count >= 1 && count <= kMaxDiis, "too many DIIS vectors");
  CC4S_ASSERT(t0 >= 0 && len >= 0 && t0 + len <= packed_set_size(No, Nv),
              "packed range out of bounds");
  DotArgs a;
  for (int k = 0; k < count; ++k) a.x[k] = x[k];
  double2 *part = static_cast<double2 *>(partials);
  double2 *res = part + static_cast<size_t>(kReduceBlocks) * count;
  if (f == Field::Complex)
    dots_packed_kernel<true><<<kReduceBlocks, kBlock, 0, s>>>(
        a, count, static_cast<const double2 *>(y), Nv * No, t0, len, part);
  else
    dots_packed_kernel<false><<<kReduceBlocks, kBlock, 0, s>>>(
        a, count, static_cast<const double *>(y), Nv * No, t0, len, part);
  CC4S_HIP(hipGetLastError());
  finish_kernel<<<1, kBlock, 0, s>>>(part, count, kReduceBlocks, res);
  CC4S_HIP(hipGetLastError());
  std::vector<double2> h(count);
  CC4S_HIP(hipMemcpyAsync(h.data(), res, count * sizeof(double2),
                          hipMemcpyDeviceToHost, s));
  CC4S_HIP(hipStreamSynchronize(s));
  out.resize(count);
  for (int k = 0; k < count; ++k) out[k] = cplx(h[k].x, h[k].y);
}

double combine_packed(hipStream_t s, Field f, const void *const *A,
                      const void *const *R, const std::vector<double> &c,
                      int64_t No, int64_t Nv, void *out, void *partials) {
  const int count = static_cast<int>(c.size());
  CC4S_ASSERT(count >= 1 && count <= kMaxDiis, "too many DIIS vectors");
  DotArgs a;
  for (int k = 0; k < count; ++k) {
    a.A[k] = A[k];
    a.R[k] = R[k];
    a.c[k] = c[k];
  }
  double2 *part = static_cast<double2 *>(partials);
  double2 *res = part + kReduceBlocks;
  if (f == Field::Complex)
    combine_packed_kernel<true><<<kReduceBlocks, kBlock, 0, s>>>(
        a, count, static_cast<double2 *>(out), No, Nv, part);
  else
    combine_packed_kernel<false><<<kReduceBlocks, kBlock, 0, s>>>(
        a, count, static_cast<double *>(out), No, Nv, part);
  CC4S_HIP(hipGetLastError());
  finish_kernel<<<1, kBlock, 0, s>>>(part, 1, kReduceBlocks, res);
  CC4S_HIP(hipGetLastError());
  double2 h;
  CC4S_HIP(hipMemcpyAsync(&h, res, sizeof(h), hipMemcpyDeviceToHost, s));
  CC4S_HIP(hipStreamSynchronize(s));
  return h.x;
}

namespace {

// per-block count of (G,p,q), p <= q, with Gamma[G,q,p] != conj Gamma[G,p,q]
__global__ __launch_bounds__(kBlock) void hermitian_kernel(const double2 *g,
                                                           int64_t NG, int64_t Np,
                                                           double2 *partials) {
  __shared__ double2 sh[kBlock];
  const int64_t n = NG * Np * Np;
  double bad = 0.0;
  for (int64_t e = blockIdx.x * int64_t(kBlock) + threadIdx.x; e < n;
       e += int64_t(gridDim.x) * kBlock) {
    const int64_t G = e % NG;
    const int64_t r = e / NG;
    const int64_t p = r % Np, q = r / Np;
    if (p > q) continue;
    const double2 a = g[e];
    const double2 b = g[G + NG * (q + Np * p)];
    if (!(b.x == a.x && b.y == -a.y)) bad += 1.0;
  }
  const double2 s = block_sum(make_double2(bad, 0.0), sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

}  // namespace

bool is_hermitian(hipStream_t s, const void *gamma, int64_t NG, int64_t Np) {
  DeviceBuffer part((kReduceBlocks + 1) * sizeof(double2));
  double2 *p = part.as<double2>();
  hermitian_kernel<<<kReduceBlocks, kBlock, 0, s>>>(
      static_cast<const double2 *>(gamma), NG, Np, p);
  CC4S_HIP(hipGetLastError());
  finish_kernel<<<1, kBlock, 0, s>>>(p, 1, kReduceBlocks, p + kReduceBlocks);
  CC4S_HIP(hipGetLastError());
  double2 h;
  CC4S_HIP(hipMemcpyAsync(&h, p + kReduceBlocks, sizeof(h), hipMemcpyDeviceToHost, s));
  CC4S_HIP(hipStreamSynchronize(s));
  return h.x == 0.0;
}

void copy(hipStream_t s, void *y, const void *x, size_t bytes) {
  CC4S_HIP(hipMemcpyAsync(y, x, bytes, hipMemcpyDeviceToDevice, s));
}

namespace {
__global__ __launch_bounds__(kBlock) void axpby_kernel(int64_t n, double a,
                                                       const double *__restrict__ x,
                                                       double b, double *__restrict__ y) {
  for (int64_t e = blockIdx.x * int64_t(kBlock) + threadIdx.x; e < n;
       e += int64_t(gridDim.x) * kBlock)
    y[e] = a * x[e] + b * y[e];
}
}  // namespace

void axpby(hipStream_t s, int64_t n, double a, const void *x, double b, void *y) {
  if (n <= 0) return;
  axpby_kernel<<<grid_for(n), kBlock, 0, s>>>(n, a, static_cast<const double *>(x), b,
                                              static_cast<double *>(y));
  CC4S_HIP(hipGetLastError());
}

}  // namespace cc4s_amd
