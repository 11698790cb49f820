// Energy adjoint kernels: energy weights from the amplitudes, FP32 operand
// copies, the FP32 complex product of the FP32 gradient, and the per-G block
// dots of the structure factor.  The FP64 gradient products run on the FP64
// MFMA contraction engine; everything here is bandwidth-bound except
// gemm_fp32 (G*No^2*Nv^2 complex MACs, small next to one CCSD iteration).
#include "adjoint.hpp"

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

// M[k + K*n], K = No*Nv; complex k = i + No*a, real k = a + Nv*i; n = b + Nv*j
template <bool CPLX>
__global__ __launch_bounds__(kBlock) void weights_fp32_kernel(Amp<CPLX> t, float2 *M) {
  const int64_t No = t.No, Nv = t.Nv, K = No * Nv;
  const int64_t n = K * K;
  for (int64_t e = blockIdx.x * int64_t(kBlock) + threadIdx.x; e < n;
       e += int64_t(gridDim.x) * kBlock) {
    const int64_t k = e % K, col = e / K;
    const int64_t b = col % Nv, j = col / Nv;
    if constexpr (CPLX) {
      const int64_t i = k % No, a = k / No;
      const double2 z = t.z(a, b, i, j);
      M[e] = make_float2(static_cast<float>(z.x), static_cast<float>(z.y));
    } else {
      const int64_t a = k % Nv, i = k / Nv;
      M[e] = make_float2(static_cast<float>(t.zs(a, b, i, j)), 0.0f);
    }
  }
}

__global__ __launch_bounds__(kBlock) void vertex_block_fp32_kernel(
    const double2 *g, int64_t NG, int64_t Np, int64_t row0, int64_t rows,
    int64_t col0, int64_t cols, float2 *out) {
  const int64_t n = NG * rows * cols;
  for (int64_t e = blockIdx.x * int64_t(kBlock) + threadIdx.x; e < n;
       e += int64_t(gridDim.x) * kBlock) {
    const int64_t G = e % NG, xy = e / NG;
    const int64_t x = xy % rows, y = xy / rows;
    const double2 v = g[G + NG * ((row0 + x) + Np * (col0 + y))];
    out[e] = make_float2(static_cast<float>(v.x), static_cast<float>(v.y));
  }
}

// 64 (G) x 64 (n) tile, K steps of 16, 256 threads with 4 x 4 complex
// accumulators each; FP32 FMAs on the vector ALUs.
constexpr int kGM = 64, kGN = 64, kGK = 16;

__global__ __launch_bounds__(256) void gemm_fp32_kernel(
    const float2 *__restrict__ A, const float2 *__restrict__ B, int64_t sbk,
    int64_t sbn, int conjB, double2 *__restrict__ C, int64_t NG, int64_t N,
    int64_t K) {
  __shared__ float2 As[kGK][kGM];
  __shared__ float2 Bs[kGK][kGN];
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  const int64_t g0 = int64_t(blockIdx.x) * kGM;
  const int64_t n0 = int64_t(blockIdx.y) * kGN;
  const float bsign = conjB ? -1.0f : 1.0f;
  float2 acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = make_float2(0.0f, 0.0f);
  for (int64_t k0 = 0; k0 < K; k0 += kGK) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int q = tid + 256 * r;
      const int m = q % kGM, k = q / kGM;
      const int64_t G = g0 + m, kk = k0 + k;
      As[k][m] = (G < NG && kk < K) ? A[G + NG * kk] : make_float2(0.0f, 0.0f);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int q = tid + 256 * r;
      int k, nn;
      if (sbk == 1) {  // read along k
        k = q % kGK;
        nn = q / kGK;
      } else {         // read along n
        nn = q % kGN;
        k = q / kGN;
      }
      const int64_t kk = k0 + k, col = n0 + nn;
      float2 b = make_float2(0.0f, 0.0f);
      if (kk < K && col < N) {
        b = B[kk * sbk + col * sbn];
        b.y *= bsign;
      }
      Bs[k][nn] = b;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kGK; ++kk) {
      float2 a[4], b[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) a[r] = As[kk][tx + 16 * r];
#pragma unroll
      for (int c = 0; c < 4; ++c) b[c] = Bs[kk][ty + 16 * c];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          acc[r][c].x = fmaf(a[r].x, b[c].x, acc[r][c].x);
          acc[r][c].x = fmaf(-a[r].y, b[c].y, acc[r][c].x);
          acc[r][c].y = fmaf(a[r].x, b[c].y, acc[r][c].y);
          acc[r][c].y = fmaf(a[r].y, b[c].x, acc[r][c].y);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int64_t G = g0 + tx + 16 * r, col = n0 + ty + 16 * c;
      if (G < NG && col < N)
        C[G + NG * col] = make_double2(acc[r][c].x, acc[r][c].y);
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

void energy_weights_fp32(hipStream_t s, Field f, const void *T, int64_t No,
                         int64_t Nv, void *M) {
  const int64_t n = Nv * Nv * No * No;
  if (f == Field::Complex) {
    Amp<true> t{static_cast<const double2 *>(T), static_cast<const double2 *>(T) + Nv * No,
                No, Nv};
    weights_fp32_kernel<true><<<grid_for(n), kBlock, 0, s>>>(t, static_cast<float2 *>(M));
  } else {
    Amp<false> t{static_cast<const double *>(T), static_cast<const double *>(T) + Nv * No,
                 No, Nv};
    weights_fp32_kernel<false><<<grid_for(n), kBlock, 0, s>>>(t, static_cast<float2 *>(M));
  }
  CC4S_HIP(hipGetLastError());
}

void vertex_block_fp32(hipStream_t s, const void *gamma, int64_t NG, int64_t Np,
                       int64_t row0, int64_t rows, int64_t col0, int64_t cols,
                       void *out) {
  CC4S_ASSERT(row0 >= 0 && col0 >= 0 && row0 + rows <= Np && col0 + cols <= Np,
              "vertex block out of range");
  vertex_block_fp32_kernel<<<grid_for(NG * rows * cols), kBlock, 0, s>>>(
      static_cast<const double2 *>(gamma), NG, Np, row0, rows, col0, cols,
      static_cast<float2 *>(out));
  CC4S_HIP(hipGetLastError());
}

void gemm_fp32(hipStream_t s, const void *A, const void *B, int64_t sbk,
               int64_t sbn, bool conjB, void *C, int64_t NG, int64_t N,
               int64_t K) {
  CC4S_ASSERT(NG > 0 && N > 0 && K > 0, "empty FP32 product");
  const int64_t gy = (N + kGN - 1) / kGN;
  CC4S_ASSERT(gy <= 65535, "FP32 product too wide");
  dim3 grid(static_cast<unsigned>((NG + kGM - 1) / kGM), static_cast<unsigned>(gy));
  gemm_fp32_kernel<<<grid, 256, 0, s>>>(static_cast<const float2 *>(A),
                                         static_cast<const float2 *>(B), sbk, sbn,
                                         conjB ? 1 : 0, static_cast<double2 *>(C), NG,
                                         N, K);
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
