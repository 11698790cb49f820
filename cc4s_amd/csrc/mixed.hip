// Split-bf16 complex contraction (the mixed-precision path).
//
// Every FP64 complex operand element x = xr + i xi is split into four bf16
// values xr = rh + rl, xi = ih + il (rh = bf16(xr), rl = bf16(xr - rh), the
// same for xi), i.e. about 16 significant bits per part.  An operand is first
// packed, through the same offset tables as the FP64 kernel, into a dense
// zero-padded [k][m] array of 8-byte (rh|rl<<16, ih|il<<16) words.  The
// multiply kernel then feeds v_mfma_f32_32x32x16_bf16 with, per lane, one
// complex k expanded to eight bf16:
//   A:    [ah  ah  al  al  aih aih ail ail]   (ah = rh of a, ...)
//   B_re: [bh  bl  bh  bl  -bih -bil -bih -bil]
//   B_im: [bih bil bih bil  bh  bl  bh  bl ]
// so that the 8-term dot products are exactly Re(a b) and Im(a b) of the split
// values, accumulated in FP32 (32x32x16 = two complex k per instruction;
// lanes 0-31 carry k, lanes 32-63 k+1).  Results are written to C in FP64
// (alpha * acc + beta * C) or to FP64 split-K partials.  K is processed in
// chunks, so FP32 sums never run over more than one chunk.
//
// FP16 variant: the same split with IEEE half parts (11 significant bits
// each, about 22 per real part).  Half has a narrow exponent range, so each
// operand is first multiplied by a power of two that brings its largest
// magnitude just below 2^15 (a max-|x| reduction over the operand), and the
// kernel multiplies the FP32 sums by the inverse of both scales; every
// scaling is exact.  Parts below 2^-24 of the scaled range flush towards
// zero, which is far below the kept precision of the large elements.
#include "contract.hpp"

#include <algorithm>

namespace cc4s_amd {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned int bf16_bits_rne(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

__device__ __forceinline__ float bf16_value(unsigned int b) {
  return __uint_as_float(b << 16);
}

// (hi | lo << 16) of one real part
__device__ __forceinline__ unsigned int split_part(double x) {
  const unsigned int h = bf16_bits_rne(static_cast<float>(x));
  const double r = x - static_cast<double>(bf16_value(h));
  const unsigned int l = bf16_bits_rne(static_cast<float>(r));
  return h | (l << 16);
}

// (hi | lo << 16) of one real part in IEEE half, x already scaled
__device__ __forceinline__ unsigned int split_part_f16(double x) {
  const _Float16 h = static_cast<_Float16>(static_cast<float>(x));
  const double r = x - static_cast<double>(static_cast<float>(h));
  const _Float16 l = static_cast<_Float16>(static_cast<float>(r));
  return static_cast<unsigned int>(__builtin_bit_cast(unsigned short, h)) |
         (static_cast<unsigned int>(__builtin_bit_cast(unsigned short, l)) << 16);
}

constexpr int kPackTile = 64;
constexpr int kMaxBlocks = 1024;

// per-block max(|re|, |im|) over the operand elements offMN[mn] + offK[k].
// Work items are segments of 1024 consecutive elements of one line of the
// contiguous index (k if kcontig, else mn); a wave takes one item, its lanes
// read 16 elements each with coalesced loads.
constexpr int kSegment = 1024;

__global__ __launch_bounds__(256) void maxabs_kernel(const double2 *X, const int64_t *offMN,
                                                     const int64_t *offK, int64_t MN,
                                                     int64_t K, bool kcontig,
                                                     double *partials) {
  __shared__ double sh[256];
  double m = 0.0;
  const int lane = threadIdx.x & 63;
  const int64_t wave = blockIdx.x * 4ll + (threadIdx.x >> 6);
  const int64_t waves = gridDim.x * 4ll;
  const int64_t lines = kcontig ? MN : K, len = kcontig ? K : MN;
  const int64_t *lineOff = kcontig ? offMN : offK;
  const int64_t *inOff = kcontig ? offK : offMN;
  const int64_t nseg = (len + kSegment - 1) / kSegment;
  for (int64_t item = wave; item < lines * nseg; item += waves) {
    const int64_t l = item / nseg;
    const int64_t i0 = (item - l * nseg) * kSegment;
    const int64_t i1 = min(len, i0 + kSegment);
    const int64_t base = lineOff[l];
#pragma unroll 4
    for (int64_t i = i0 + lane; i < i1; i += 64) {
      const double2 x = X[base + inOff[i]];
      m = fmax(m, fmax(fabs(x.x), fabs(x.y)));
    }
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

// scale[0] = 2^(15 - e) for max = f 2^e, f in [0.5, 1): the scaled maximum
// lies in [2^14, 2^15); an all-zero operand gets scale 1
__global__ __launch_bounds__(256) void scale_kernel(const double *partials, int nblocks,
                                                    double *scale) {
  __shared__ double sh[256];
  double m = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += 256) m = fmax(m, partials[b]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    int e = 0;
    (void)frexp(sh[0], &e);
    scale[0] = sh[0] > 0.0 ? ldexp(1.0, 15 - e) : 1.0;
  }
}

struct PackArgs {
  const double2 *X;
  const int64_t *offMN;  // offset of each row (m or n) index
  const int64_t *offK;   // offset of each k
  int64_t MN, K;         // logical extents
  int64_t k0;            // first k of this chunk
  int64_t Kc;            // k rows written (padded chunk length)
  int64_t MNp;           // padded row length of the packed array
  double sign;           // -1 conjugates
  bool kcontig;          // source contiguous along k
  uint2 *out;            // [Kc][MNp]
  const double *scale;   // FP16: the operand's power-of-two scale
};

// packs a 64(k) x 64(mn) tile: reads in the source's contiguous order,
// transposes through LDS, writes mn-fastest
template <bool FP16>
__global__ __launch_bounds__(256) void pack_kernel(PackArgs p) {
  __shared__ uint2 tile[kPackTile][kPackTile + 1];
  // the tile's offsets, read once (-1: outside the operand)
  __shared__ int64_t oK[kPackTile], oMN[kPackTile];
  const double sc = FP16 ? p.scale[0] : 1.0;
  const int64_t mnb = static_cast<int64_t>(blockIdx.x) * kPackTile;
  const int64_t kb = static_cast<int64_t>(blockIdx.y) * kPackTile;
  const int t = threadIdx.x;
  if (t < kPackTile) {
    const int64_t k = p.k0 + kb + t;
    oK[t] = (kb + t < p.Kc && k < p.K) ? p.offK[k] : -1;
  } else if (t < 2 * kPackTile) {
    const int64_t mn = mnb + (t - kPackTile);
    oMN[t - kPackTile] = mn < p.MN ? p.offMN[mn] : -1;
  }
  __syncthreads();
  for (int e = t; e < kPackTile * kPackTile; e += 256) {
    const int a = e % kPackTile, b = e / kPackTile;
    const int kk = p.kcontig ? a : b, mm = p.kcontig ? b : a;
    uint2 v = make_uint2(0u, 0u);
    const int64_t ok = oK[kk], om = oMN[mm];
    if (ok >= 0 && om >= 0) {
      const double2 x = p.X[om + ok];
      if constexpr (FP16) {
        v.x = split_part_f16(sc * x.x);
        v.y = split_part_f16(sc * p.sign * x.y);
      } else {
        v.x = split_part(x.x);
        v.y = split_part(p.sign * x.y);
      }
    }
    tile[kk][mm] = v;
  }
  __syncthreads();
  for (int e = t; e < kPackTile * kPackTile; e += 256) {
    const int mm = e % kPackTile, kk = e / kPackTile;
    const int64_t krow = kb + kk, mn = mnb + mm;
    if (krow < p.Kc && mn < p.MNp) p.out[krow * p.MNp + mn] = tile[kk][mm];
  }
}

struct BgArgs {
  const uint2 *A;  // [Kc][Mp]
  const uint2 *B;  // [Kc][Np]
  double *C;
  double *W;
  const int64_t *rowC, *colC;
  int64_t M, N, Mp, Np, Kc, kChunk;
  double alpha, beta;
  int tilesM, tilesN;
  const double *scaleA, *scaleB;  // FP16: the operands' scales
};

// the eight 16-bit parts of one lane's operand (bf16 or half bit patterns);
// the sign flip is bit 15 of each part in both formats
__device__ __forceinline__ u32x4 expand_a(uint2 c) {
  u32x4 w;
  w[0] = (c.x & 0xffffu) * 0x10001u;
  w[1] = (c.x >> 16) * 0x10001u;
  w[2] = (c.y & 0xffffu) * 0x10001u;
  w[3] = (c.y >> 16) * 0x10001u;
  return w;
}
__device__ __forceinline__ u32x4 expand_b_re(uint2 c) {
  const unsigned int ni = c.y ^ 0x80008000u;
  u32x4 w = {c.x, c.x, ni, ni};
  return w;
}
__device__ __forceinline__ u32x4 expand_b_im(uint2 c) {
  u32x4 w = {c.y, c.y, c.x, c.x};
  return w;
}

template <bool FP16>
__device__ __forceinline__ f32x16 mfma_split(u32x4 a, u32x4 b, f32x16 c) {
  if constexpr (FP16)
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a),
                                                  __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

constexpr int kBThreads = 512;

// 8 waves in a 2 x 4 grid over a BM x BN complex tile; each wave owns
// WM x WN = (WM/32) x (WN/32) blocks of 32x32, re and im accumulators.
template <int BM, int BN, int BK, int WM, int WN, bool FP16>
__global__ __launch_bounds__(kBThreads, 1) void bgemm_kernel(BgArgs p) {
  constexpr int MB = WM / 32, NB = WN / 32;
  constexpr int WAVES_N = BN / WN;
  static_assert((BM / WM) * (BN / WN) * 64 == kBThreads, "eight waves");
  static_assert(BK % 2 == 0, "two complex k per MFMA");
  constexpr int QA = BK * BM / 2 / kBThreads;  // uint4 (two elements) per thread
  constexpr int QB = BK * BN / 2 / kBThreads;
  static_assert(QA >= 1 && QB >= 1, "tile too small");

  __shared__ uint2 As[2][BK][BM];
  __shared__ uint2 Bs[2][BK][BN];

  const int t = threadIdx.x;
  const int ntiles = p.tilesM * p.tilesN;
  int g = blockIdx.x;
  if ((ntiles & 7) == 0) g = (g & 7) * (ntiles >> 3) + (g >> 3);
  constexpr int GM = 8;
  const int per_group = GM * p.tilesN;
  const int first_m = (g / per_group) * GM;
  const int gsize = min(p.tilesM - first_m, GM);
  const int tm = first_m + (g % per_group) % gsize;
  const int tn = (g % per_group) / gsize;
  const int64_t m0 = static_cast<int64_t>(tm) * BM;
  const int64_t n0 = static_cast<int64_t>(tn) * BN;
  const int64_t kbeg = static_cast<int64_t>(blockIdx.y) * p.kChunk;
  const int64_t kend = min(p.Kc, kbeg + p.kChunk);
  const int nstages = kend > kbeg ? static_cast<int>((kend - kbeg) / BK) : 0;

  // staging registers, named so that none of them is placed in LDS
  static_assert(QA == 2 && QB == 4, "staging is written out for 2 + 4 words");
  uint4 ra0, ra1, rb0, rb1, rb2, rb3;
  const int akk = t / (BM / 2), amm = 2 * (t % (BM / 2));  // + q * (kBThreads / (BM/2)) rows
  const int bkk = t / (BN / 2), bnn = 2 * (t % (BN / 2));
  constexpr int AR = kBThreads / (BM / 2), BR = kBThreads / (BN / 2);
  auto load = [&](int s) {
    const int64_t kb = kbeg + static_cast<int64_t>(s) * BK;
    const uint2 *a = p.A + (kb + akk) * p.Mp + m0 + amm;
    const uint2 *b = p.B + (kb + bkk) * p.Np + n0 + bnn;
    ra0 = *reinterpret_cast<const uint4 *>(a);
    ra1 = *reinterpret_cast<const uint4 *>(a + AR * p.Mp);
    rb0 = *reinterpret_cast<const uint4 *>(b);
    rb1 = *reinterpret_cast<const uint4 *>(b + BR * p.Np);
    rb2 = *reinterpret_cast<const uint4 *>(b + 2 * BR * p.Np);
    rb3 = *reinterpret_cast<const uint4 *>(b + 3 * BR * p.Np);
  };
  auto store = [&](int buf) {
    *reinterpret_cast<uint4 *>(&As[buf][akk][amm]) = ra0;
    *reinterpret_cast<uint4 *>(&As[buf][akk + AR][amm]) = ra1;
    *reinterpret_cast<uint4 *>(&Bs[buf][bkk][bnn]) = rb0;
    *reinterpret_cast<uint4 *>(&Bs[buf][bkk + BR][bnn]) = rb1;
    *reinterpret_cast<uint4 *>(&Bs[buf][bkk + 2 * BR][bnn]) = rb2;
    *reinterpret_cast<uint4 *>(&Bs[buf][bkk + 3 * BR][bnn]) = rb3;
  };

  const int lane = t & 63;
  const int w = t >> 6;
  const int wm = w / WAVES_N, wn = w % WAVES_N;
  f32x16 accR[MB][NB], accI[MB][NB];
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        accR[i][j][r] = 0.f;
        accI[i][j][r] = 0.f;
      }
    }

  if (nstages > 0) {
    load(0);
    store(0);
    if (nstages > 1) load(1);
    __syncthreads();
  }
  for (int s = 0; s < nstages; ++s) {
    const int buf = s & 1;
    static_assert(MB == 2 && NB == 2, "fragments are written out for 2 x 2 blocks");
#pragma unroll
    for (int kp = 0; kp < BK / 2; ++kp) {
      const int kk = 2 * kp + (lane >> 5);
      const int ra0 = wm * WM + (lane & 31), rb0 = wn * WN + (lane & 31);
      // named registers (not arrays), so no fragment is ever placed in LDS
      const u32x4 a0 = expand_a(As[buf][kk][ra0]);
      const u32x4 a1 = expand_a(As[buf][kk][ra0 + 32]);
      const uint2 c0 = Bs[buf][kk][rb0];
      const uint2 c1 = Bs[buf][kk][rb0 + 32];
      const u32x4 br0 = expand_b_re(c0), bi0 = expand_b_im(c0);
      const u32x4 br1 = expand_b_re(c1), bi1 = expand_b_im(c1);
      accR[0][0] = mfma_split<FP16>(a0, br0, accR[0][0]);
      accI[0][0] = mfma_split<FP16>(a0, bi0, accI[0][0]);
      accR[0][1] = mfma_split<FP16>(a0, br1, accR[0][1]);
      accI[0][1] = mfma_split<FP16>(a0, bi1, accI[0][1]);
      accR[1][0] = mfma_split<FP16>(a1, br0, accR[1][0]);
      accI[1][0] = mfma_split<FP16>(a1, bi0, accI[1][0]);
      accR[1][1] = mfma_split<FP16>(a1, br1, accR[1][1]);
      accI[1][1] = mfma_split<FP16>(a1, bi1, accI[1][1]);
    }
    if (s + 1 < nstages) {
      store(buf ^ 1);
      if (s + 2 < nstages) load(s + 2);
    }
    __syncthreads();
  }

  // lane l, register r of a 32x32 block: row 8*(r/4) + 4*(l/32) + r%4, column l%32
  const double unscale = FP16 ? 1.0 / (p.scaleA[0] * p.scaleB[0]) : 1.0;
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const int64_t n = n0 + wn * WN + j * 32 + (lane & 31);
    if (n >= p.N) continue;
    const int64_t cc = p.W ? 0 : p.colC[n];
#pragma unroll
    for (int i = 0; i < MB; ++i) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t m = m0 + wm * WM + i * 32 + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
        if (m >= p.M) continue;
        const double vr = unscale * static_cast<double>(accR[i][j][r]);
        const double vi = unscale * static_cast<double>(accI[i][j][r]);
        if (p.W) {
          const int64_t o = (static_cast<int64_t>(blockIdx.y) * p.M + m) * p.N + n;
          reinterpret_cast<double2 *>(p.W)[o] = make_double2(vr, vi);
        } else {
          double2 *c = reinterpret_cast<double2 *>(p.C) + p.rowC[m] + cc;
          double2 out = make_double2(p.alpha * vr, p.alpha * vi);
          if (p.beta != 0.0) {
            const double2 old = *c;
            out.x += p.beta * old.x;
            out.y += p.beta * old.y;
          }
          *c = out;
        }
      }
    }
  }
}

constexpr int kXBM = 128, kXBN = 256, kXBK = 16, kXWM = 64, kXWN = 64;

}  // namespace

MixedTile mixed_tile() { return MixedTile{kXBM, kXBN, kXBK}; }

void launch_split_scale(const double *X, const int64_t *offMN, const int64_t *offK,
                        int64_t MN, int64_t K, bool kcontig, double *partials,
                        double *scale, hipStream_t s) {
  const int64_t n = MN * K;
  const int blocks = static_cast<int>(std::max<int64_t>(
      1, std::min<int64_t>(kMaxBlocks, (n + 255) / 256)));
  maxabs_kernel<<<blocks, 256, 0, s>>>(reinterpret_cast<const double2 *>(X), offMN, offK,
                                       MN, K, kcontig, partials);
  CC4S_HIP(hipGetLastError());
  scale_kernel<<<1, 256, 0, s>>>(partials, blocks, scale);
  CC4S_HIP(hipGetLastError());
}

size_t split_scale_partials_bytes() { return kMaxBlocks * sizeof(double); }

void launch_pack(const double *X, const int64_t *offMN, const int64_t *offK,
                 int64_t MN, int64_t K, int64_t k0, int64_t Kc, int64_t MNp,
                 bool conj, bool kcontig, void *out, const double *scale,
                 hipStream_t s) {
  PackArgs a;
  a.X = reinterpret_cast<const double2 *>(X);
  a.offMN = offMN;
  a.offK = offK;
  a.MN = MN;
  a.K = K;
  a.k0 = k0;
  a.Kc = Kc;
  a.MNp = MNp;
  a.sign = conj ? -1.0 : 1.0;
  a.kcontig = kcontig;
  a.out = static_cast<uint2 *>(out);
  a.scale = scale;
  dim3 grid(static_cast<unsigned>((MNp + kPackTile - 1) / kPackTile),
            static_cast<unsigned>((Kc + kPackTile - 1) / kPackTile));
  if (scale)
    pack_kernel<true><<<grid, 256, 0, s>>>(a);
  else
    pack_kernel<false><<<grid, 256, 0, s>>>(a);
  CC4S_HIP(hipGetLastError());
}

void launch_bgemm(const void *Apack, const void *Bpack, double *C, double *W,
                  const int64_t *rowC, const int64_t *colC, int64_t M, int64_t N,
                  int64_t Mp, int64_t Np, int64_t Kc, int splits, double alpha,
                  double beta, const double *scaleA, const double *scaleB,
                  hipStream_t s) {
  CC4S_ASSERT((scaleA == nullptr) == (scaleB == nullptr),
              "split-half operands need both scales");
  BgArgs a;
  a.scaleA = scaleA;
  a.scaleB = scaleB;
  a.A = static_cast<const uint2 *>(Apack);
  a.B = static_cast<const uint2 *>(Bpack);
  a.C = C;
  a.W = splits > 1 ? W : nullptr;
  a.rowC = rowC;
  a.colC = colC;
  a.M = M; a.N = N; a.Mp = Mp; a.Np = Np; a.Kc = Kc;
  const int64_t stages = Kc / kXBK;
  a.kChunk = ((stages + splits - 1) / splits) * kXBK;
  a.alpha = alpha;
  a.beta = beta;
  a.tilesM = static_cast<int>(Mp / kXBM);
  a.tilesN = static_cast<int>(Np / kXBN);
  dim3 grid(a.tilesM * a.tilesN, splits);
  if (scaleA)
    bgemm_kernel<kXBM, kXBN, kXBK, kXWM, kXWN, true><<<grid, kBThreads, 0, s>>>(a);
  else
    bgemm_kernel<kXBM, kXBN, kXBK, kXWM, kXWN, false><<<grid, kBThreads, 0, s>>>(a);
  CC4S_HIP(hipGetLastError());
}

}  // namespace cc4s_amd
