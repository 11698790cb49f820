// Energy adjoint kernels: the gradient products with the energy weights
// recomputed from the amplitudes inside the product (FP64 or FP32 MFMA,
// G*No^2*Nv^2 complex MACs per block, small next to one CCSD iteration), the
// energy weights as a tensor (for the structure factor of the finite-size
// correction), and the per-G block dots of the structure factor.
#include "adjoint.hpp"

#include <type_traits>

namespace cc4s_amd {

namespace {

constexpr int kBlock = 256;
constexpr int kDotChunks = 64;  // fixed: deterministic reduction order

inline int grid_for(int64_t n) {
  return static_cast<int>(std::min<int64_t>((n + kBlock - 1) / kBlock, 256 * 32));
}

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// tau_abij = T2_abij + T1_ai T1_bj (CoupledCluster.cxx:223-231)
template <bool CPLX>
struct Amp;
template <>
struct Amp<true> {
  const double2 *T1, *T2;
  int64_t No, Nv;
  __device__ double2 tau(int64_t a, int64_t b, int64_t i, int64_t j) const {
    const double2 t = T2[a + Nv * (b + Nv * (i + No * j))];
    const double2 s = cmul(T1[a + Nv * i], T1[b + Nv * j]);
    return make_double2(t.x + s.x, t.y + s.y);
  }
  // Z_abij = 2 tau_abij - tau_baij
  __device__ double2 z(int64_t a, int64_t b, int64_t i, int64_t j) const {
    const double2 p = tau(a, b, i, j), q = tau(b, a, i, j);
    return make_double2(2.0 * p.x - q.x, 2.0 * p.y - q.y);
  }
};
template <>
struct Amp<false> {
  const double *T1, *T2;
  int64_t No, Nv;
  __device__ double tau(int64_t a, int64_t b, int64_t i, int64_t j) const {
    return T2[a + Nv * (b + Nv * (i + No * j))] + T1[a + Nv * i] * T1[b + Nv * j];
  }
  __device__ double z(int64_t a, int64_t b, int64_t i, int64_t j) const {
    return 2.0 * tau(a, b, i, j) - tau(b, a, i, j);
  }
  // weight of the real vertex: Z_abij + Z_baji
  __device__ double zs(int64_t a, int64_t b, int64_t i, int64_t j) const {
    return z(a, b, i, j) + z(b, a, j, i);
  }
};

template <bool CPLX>
__global__ __launch_bounds__(kBlock) void weights_kernel(Amp<CPLX> t, void *Zout) {
  const int64_t No = t.No, Nv = t.Nv;
  const int64_t n = Nv * Nv * No * No;
  for (int64_t e = blockIdx.x * int64_t(kBlock) + threadIdx.x; e < n;
       e += int64_t(gridDim.x) * kBlock) {
    int64_t r = e;
    const int64_t a = r % Nv; r /= Nv;
    const int64_t b = r % Nv; r /= Nv;
    const int64_t i = r % No;
    const int64_t j = r / No;
    if constexpr (CPLX)
      static_cast<double2 *>(Zout)[e] = t.z(a, b, i, j);
    else
      static_cast<double *>(Zout)[e] = t.zs(a, b, i, j);
  }
}

// ---------------------------------------------------------------------------
// Fused gradient products: C[G, n] = sum_k Gamma_blk[G, k] W(k, n), with the
// energy weights W recomputed from T1 and T2 for every tile as it is staged
// into LDS (no Z tensor exists).  Tile 128 (G) x 64 (n), 16 k per stage,
// double-buffered LDS, 8 waves of 32 x 32 on 16x16x4 MFMA; FP64 uses the
// three-product complex multiply, FP32 the four-product one (FP32 sums
// cancel too much in P3 - P1 - P2).  C is complex FP64, G fastest.
//   kHP (complex): k = (b,j), n = (a,i), A = Gamma[G, No+b, j],
//                  W = conj Z_abij, C at G + NG (i + No a)
//   kPH (complex): k = (a,i), n = (b,j), A = Gamma[G, i, No+a],
//                  W = Z_abij, C at G + NG (b + Nv j)
//   kRE (real):    k = (b,j), n = (a,i), A = Gamma[G, No+b, j] (Re and Im are
//                  the two real vertex components), W = Z_abij + Z_baji (real),
//                  C at G + NG (a + Nv i)
enum { kHP = 0, kPH = 1, kRE = 2 };

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  using acc = dbl4;
  using pair = double2;
  __device__ static acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // row of register r of the 16x16 block held by `lane`
  __device__ static int row(int r, int lane) { return 4 * r + (lane >> 4); }
  __device__ static pair make(double x, double y) { return make_double2(x, y); }
};
template <>
struct Mfma<float> {
  using acc = flt4;
  using pair = float2;
  __device__ static acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int r, int lane) { return 4 * (lane >> 4) + r; }
  __device__ static pair make(double x, double y) {
    return make_float2(static_cast<float>(x), static_cast<float>(y));
  }
};

struct GradArgs {
  const double2 *gamma;  // [NG, Np, Np]
  const double2 *T1c, *T2c;  // complex amplitudes
  const double *T1r, *T2r;   // real amplitudes
  double2 *C;
  int64_t NG, Np, No, Nv;
};

constexpr int kAT = 512;                  // 8 waves
constexpr int kABM = 128, kABN = 64, kABK = 16;
constexpr int kALA = kABM + 1, kALW = kABN + 1;

template <typename T, int MODE>
__global__ __launch_bounds__(kAT, 1) void grad_kernel(GradArgs p) {
  using M = Mfma<T>;
  using P = typename M::pair;
  using Acc = typename M::acc;
  constexpr bool M3 = std::is_same<T, double>::value && MODE != kRE;
  __shared__ P As[2][kABK][kALA];
  __shared__ P Ws[2][kABK][kALW];

  const int64_t NG = p.NG, No = p.No, Nv = p.Nv, Np = p.Np;
  const int64_t K = No * Nv, N = No * Nv;
  const uint32_t nv32 = static_cast<uint32_t>(Nv);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int64_t g0 = static_cast<int64_t>(blockIdx.x) * kABM;
  const int64_t n0 = static_cast<int64_t>(blockIdx.y) * kABN;
  const int nst = static_cast<int>((K + kABK - 1) / kABK);

  // A elements of this thread: G = g0 + t % 128, k = 4e + t / 128 of a stage
  const int ag = t % kABM, ak = t / kABM;
  const bool gOk = g0 + ag < NG;
  const int64_t gIdx = gOk ? g0 + ag : 0;
  // W elements: n = n0 + t % 64, k = 8e + t / 64 of a stage
  const int wn = t % kABN, wk = t / kABN;
  const int64_t nn = n0 + wn;
  const bool nOk = nn < N;
  // the n index pair: (a, i) for kHP and kRE, (b, j) for kPH; fastest first
  const int64_t nf = nOk ? nn % Nv : 0, ns = nOk ? nn / Nv : 0;

  double2 ra[4];
  bool kaOk[4];
  // amplitude values of the two W elements, combined at the LDS store
  double2 t2a[2], t2b[2], t1a[2], t1b[2], t1c[2], t1d[2];
  double r2a[2], r2b[2], r2c[2], r2d[2], r1a[2], r1b[2], r1c[2], r1d[2];
  bool kwOk[2];

  auto t2c = [&](int64_t a, int64_t b, int64_t i, int64_t j) {
    return p.T2c[a + Nv * (b + Nv * (i + No * j))];
  };
  auto t2r = [&](int64_t a, int64_t b, int64_t i, int64_t j) {
    return p.T2r[a + Nv * (b + Nv * (i + No * j))];
  };
  auto load = [&](int s) {
    const int64_t kb = static_cast<int64_t>(s) * kABK;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t k = kb + 4 * e + ak;
      kaOk[e] = k < K;
      const int64_t kk = kaOk[e] ? k : 0;
      // (b, j) or (a, i); 32-bit division (No Nv < 2^31)
      const int64_t kf = static_cast<uint32_t>(kk) % nv32, ks = static_cast<uint32_t>(kk) / nv32;
      // Gamma[G, P, Q] at G + NG (P + Np Q)
      const int64_t P = MODE == kPH ? ks : No + kf, Q = MODE == kPH ? No + kf : ks;
      ra[e] = p.gamma[gIdx + NG * (P + Np * Q)];
    }
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int64_t k = kb + 8 * e + wk;
      kwOk[e] = k < K && nOk;
      const int64_t kk = k < K ? k : 0;
      const int64_t kf = static_cast<uint32_t>(kk) % nv32, ks = static_cast<uint32_t>(kk) / nv32;
      int64_t a, b, i, j;
      if (MODE == kPH) {
        a = kf; i = ks; b = nf; j = ns;
      } else {
        b = kf; j = ks; a = nf; i = ns;
      }
      if constexpr (MODE != kRE) {
        t2a[e] = t2c(a, b, i, j);
        t2b[e] = t2c(b, a, i, j);
        t1a[e] = p.T1c[a + Nv * i];
        t1b[e] = p.T1c[b + Nv * j];
        t1c[e] = p.T1c[b + Nv * i];
        t1d[e] = p.T1c[a + Nv * j];
      } else {
        r2a[e] = t2r(a, b, i, j);
        r2b[e] = t2r(b, a, j, i);
        r2c[e] = t2r(b, a, i, j);
        r2d[e] = t2r(a, b, j, i);
        r1a[e] = p.T1r[a + Nv * i];
        r1b[e] = p.T1r[b + Nv * j];
        r1c[e] = p.T1r[b + Nv * i];
        r1d[e] = p.T1r[a + Nv * j];
      }
    }
  };
  auto cmul = [](double2 x, double2 y) {
    return make_double2(x.x * y.x - x.y * y.y, x.x * y.y + x.y * y.x);
  };
  // out-of-range elements are zeroed here, component by component (a select
  // of whole pairs would go through scratch memory)
  auto store = [&](int buf) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const bool ok = gOk && kaOk[e];
      As[buf][4 * e + ak][ag] = M::make(ok ? ra[e].x : 0.0, ok ? ra[e].y : 0.0);
    }
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      double zr, zi;
      if constexpr (MODE != kRE) {
        // Z_abij = 2 tau_abij - tau_baij, tau = T2 + T1 T1 (CoupledCluster.cxx:223-231)
        const double2 p1 = cmul(t1a[e], t1b[e]), p2 = cmul(t1c[e], t1d[e]);
        zr = 2.0 * (t2a[e].x + p1.x) - (t2b[e].x + p2.x);
        zi = 2.0 * (t2a[e].y + p1.y) - (t2b[e].y + p2.y);
        if (MODE == kHP) zi = -zi;
      } else {
        // Z_abij + Z_baji
        const double p1 = r1a[e] * r1b[e], p2 = r1c[e] * r1d[e];
        zr = 2.0 * (r2a[e] + p1) - (r2c[e] + p2) + 2.0 * (r2b[e] + p1) - (r2d[e] + p2);
        zi = 0.0;
      }
      Ws[buf][8 * e + wk][wn] = M::make(kwOk[e] ? zr : 0.0, kwOk[e] ? zi : 0.0);
    }
  };

  const int wr = w / 2, wc = w % 2;  // the wave's 32 x 32 part of the tile
  Acc a1[2][2], a2[2][2], a3[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      a1[i][j] = Acc{0, 0, 0, 0};
      a2[i][j] = Acc{0, 0, 0, 0};
      a3[i][j] = Acc{0, 0, 0, 0};
    }
  load(0);
  store(0);
  __syncthreads();
  for (int s = 0; s < nst; ++s) {
    const int buf = s & 1;
    if (s + 1 < nst) load(s + 1);
#pragma unroll
    for (int kq = 0; kq < kABK / 4; ++kq) {
      const int kr = kq * 4 + (lane >> 4);
      P x[2], y[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) x[i] = As[buf][kr][wr * 32 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; ++j) y[j] = Ws[buf][kr][wc * 32 + j * 16 + (lane & 15)];
      if constexpr (MODE == kRE) {
        // real W: re and im of Gamma times W
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            a1[i][j] = M::mma(x[i].x, y[j].x, a1[i][j]);
            a2[i][j] = M::mma(x[i].y, y[j].x, a2[i][j]);
          }
      } else if constexpr (M3) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) a1[i][j] = M::mma(x[i].x, y[j].x, a1[i][j]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) a2[i][j] = M::mma(x[i].y, y[j].y, a2[i][j]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            a3[i][j] = M::mma(x[i].x + x[i].y, y[j].x + y[j].y, a3[i][j]);
      } else {
        // four products: a1 = re, a2 = im
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            a1[i][j] = M::mma(x[i].x, y[j].x, a1[i][j]);
            a2[i][j] = M::mma(x[i].x, y[j].y, a2[i][j]);
          }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            a1[i][j] = M::mma(-x[i].y, y[j].y, a1[i][j]);
            a2[i][j] = M::mma(x[i].y, y[j].x, a2[i][j]);
          }
      }
    }
    // the other buffer was read in stage s-1, released by its barrier
    if (s + 1 < nst) store(buf ^ 1);
    __syncthreads();
  }

#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int64_t n = n0 + wc * 32 + j * 16 + (lane & 15);
    if (n >= N) continue;
    const int64_t f = n % Nv, sl = n / Nv;
    // output column of n: (i + No a) for kHP, (b + Nv j) for kPH, (a + Nv i) for kRE
    const int64_t col = MODE == kHP ? sl + No * f : n;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t G = g0 + wr * 32 + i * 16 + M::row(r, lane);
        if (G >= NG) continue;
        double re, im;
        if constexpr (M3) {
          const double p1 = a1[i][j][r], p2 = a2[i][j][r];
          re = p1 - p2;
          im = a3[i][j][r] - (p1 + p2);
        } else {
          re = a1[i][j][r];
          im = a2[i][j][r];
        }
        p.C[G + NG * col] = make_double2(re, im);
      }
  }
}

// partials[c * NG + G]: chunk c of the flattened (x, y) range of one block
__global__ __launch_bounds__(kBlock) void block_dots_kernel(
    const double2 *gamma, int64_t NG, int64_t Np, int64_t row0, int64_t rows,
    int64_t col0, int64_t cols, const double2 *g, double *partials) {
  const int64_t G = blockIdx.x * int64_t(kBlock) + threadIdx.x;
  if (G >= NG) return;
  const int64_t n = rows * cols;
  const int64_t c = blockIdx.y;
  const int64_t e0 = n * c / kDotChunks, e1 = n * (c + 1) / kDotChunks;
  double s = 0.0;
  for (int64_t e = e0; e < e1; ++e) {
    const int64_t x = e % rows, y = e / rows;
    const double2 v = gamma[G + NG * ((row0 + x) + Np * (col0 + y))];
    const double2 w = g[G + NG * e];
    s += v.x * w.x + v.y * w.y;  // Re(conj(v) w)
  }
  partials[c * NG + G] = s;
}

__global__ __launch_bounds__(kBlock) void block_dots_finish_kernel(
    const double *partials, int64_t NG, double factor, int accumulate, double *out) {
  const int64_t G = blockIdx.x * int64_t(kBlock) + threadIdx.x;
  if (G >= NG) return;
  double s = 0.0;
  for (int c = 0; c < kDotChunks; ++c) s += partials[c * NG + G];
  out[G] = (accumulate ? out[G] : 0.0) + factor * s;
}

}  // namespace

void energy_weights(hipStream_t s, Field f, const void *T, int64_t No,
                    int64_t Nv, void *Z) {
  const int64_t n = Nv * Nv * No * No;
  if (f == Field::Complex) {
    Amp<true> t{static_cast<const double2 *>(T), static_cast<const double2 *>(T) + Nv * No,
                No, Nv};
    weights_kernel<true><<<grid_for(n), kBlock, 0, s>>>(t, Z);
  } else {
    Amp<false> t{static_cast<const double *>(T), static_cast<const double *>(T) + Nv * No,
                 No, Nv};
    weights_kernel<false><<<grid_for(n), kBlock, 0, s>>>(t, Z);
  }
  CC4S_HIP(hipGetLastError());
}

void energy_gradient_product(hipStream_t s, Field f, bool fp32, GradientBlock blk,
                             const void *gamma, const void *T, int64_t NG, int64_t No,
                             int64_t Nv, void *C) {
  CC4S_ASSERT(NG > 0 && No > 0 && Nv > 0, "empty gradient product");
  CC4S_ASSERT(f == Field::Complex || blk == GradientBlock::PH,
              "the real field has the ph gradient block only");
  GradArgs a{};
  a.gamma = static_cast<const double2 *>(gamma);
  a.C = static_cast<double2 *>(C);
  a.NG = NG;
  a.No = No;
  a.Nv = Nv;
  a.Np = No + Nv;
  if (f == Field::Complex) {
    a.T1c = static_cast<const double2 *>(T);
    a.T2c = a.T1c + Nv * No;
  } else {
    a.T1r = static_cast<const double *>(T);
    a.T2r = a.T1r + Nv * No;
  }
  const int64_t gy = (No * Nv + kABN - 1) / kABN;
  CC4S_ASSERT(gy <= 65535 && No * Nv < (int64_t(1) << 31), "gradient product too wide");
  const dim3 grid(static_cast<unsigned>((NG + kABM - 1) / kABM), static_cast<unsigned>(gy));
  const int mode = f == Field::Real ? kRE : (blk == GradientBlock::HP ? kHP : kPH);
#define CC4S_GRAD(T_, M_) grad_kernel<T_, M_><<<grid, dim3(kAT), 0, s>>>(a)
  if (fp32) {
    if (mode == kHP) CC4S_GRAD(float, kHP);
    else if (mode == kPH) CC4S_GRAD(float, kPH);
    else CC4S_GRAD(float, kRE);
  } else {
    if (mode == kHP) CC4S_GRAD(double, kHP);
    else if (mode == kPH) CC4S_GRAD(double, kPH);
    else CC4S_GRAD(double, kRE);
  }
#undef CC4S_GRAD
  CC4S_HIP(hipGetLastError());
}

size_t block_dots_partials_size(int64_t NG) {
  return static_cast<size_t>(kDotChunks) * NG * sizeof(double);
}

void block_dots(hipStream_t s, const void *gamma, int64_t NG, int64_t Np,
                int64_t row0, int64_t rows, int64_t col0, int64_t cols,
                const void *g, double factor, bool accumulate, void *partials,
                double *out) {
  CC4S_ASSERT(row0 >= 0 && col0 >= 0 && row0 + rows <= Np && col0 + cols <= Np,
              "vertex block out of range");
  const unsigned gx = static_cast<unsigned>((NG + kBlock - 1) / kBlock);
  block_dots_kernel<<<dim3(gx, kDotChunks), kBlock, 0, s>>>(
      static_cast<const double2 *>(gamma), NG, Np, row0, rows, col0, cols,
      static_cast<const double2 *>(g), static_cast<double *>(partials));
  CC4S_HIP(hipGetLastError());
  block_dots_finish_kernel<<<gx, kBlock, 0, s>>>(static_cast<const double *>(partials),
                                                 NG, factor, accumulate ? 1 : 0, out);
  CC4S_HIP(hipGetLastError());
}

}  // namespace cc4s_amd
