// Split-16-bit complex contraction (the mixed-precision path).
//
// Every FP64 complex operand element x = xr + i xi is split into four 16-bit
// values xr = rh + rl, xi = ih + il (rh = h(xr), rl = h(xr - rh), the same
// for xi; h rounds to bf16 or, with a per-operand power-of-two scale, to
// IEEE half).  An operand is first packed, through the same offset tables as
// the FP64 kernel, into a dense zero-padded array of 8-byte words
// (rh|rl<<16, ih|il<<16), two consecutive k per 16-byte pair.
//
// The multiply kernel feeds v_mfma_f32_32x32x16_{f16,bf16}, whose A and B
// operands carry 8 values per lane (lanes 0-31: k-slots 0-7 of their row or
// column, lanes 32-63: slots 8-15).  Default (12 products per complex
// multiply-add): the six products that keep about 22 (half) or 16 (bf16)
// significant bits of each real part,
//   Re(a b) = ah bh + ah bl + al bh - aih bih - aih bil - ail bih,
//   Im(a b) = ah bih + ah bil + al bih + aih bh + aih bl + ail bh,
// are laid along k as A = [ah ah al aih aih ail], B_re = [bh bl bh -bih -bil
// -bih], B_im = [bih bil bih bh bl bh]; four complex k of a half-wave fill
// 24 slots, i.e. three MFMAs for Re and three for Im per 8 complex k.  The
// dropped lo*lo products lie 2^-22 (half) below the product, under the
// rounding of the FP32 sums.  With 16 products (terms = 16) one complex k
// fills the 8 slots of a lane, A = [ah ah al al aih aih ail ail],
// B_re = [bh bl bh bl -bih -bil -bih -bil], B_im = [bih bil bih bil bh bl bh
// bl], two complex k per MFMA.  Sums are FP32; results go to C in FP64
// (alpha * acc + beta * C) or to FP64 split-K partials.  K is processed in
// chunks, so FP32 sums never run over more than one chunk.
//
// Half parts have a narrow exponent range, so each operand is first
// multiplied by a power of two that brings its largest magnitude just below
// 2^15 (a max-|x| reduction over the operand), and the kernel multiplies the
// FP32 sums by the inverse of both scales; every scaling is exact.  Parts
// below 2^-24 of the scaled range flush towards zero, which is far below the
// kept precision of the large elements.
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

// per-block max(|re|, |im|) (or max(|re| + |im|) with `sumabs`, the range of
// re + im for the three-product split) over the operand elements
// offMN[mn] + offK[k].
// Work items are segments of 1024 consecutive elements of one line of the
// contiguous index (k if kcontig, else mn); a wave takes one item, its lanes
// read 16 elements each with coalesced loads.
constexpr int kSegment = 1024;

__global__ __launch_bounds__(256) void maxabs_kernel(const double2 *X, const int64_t *offMN,
                                                     const int64_t *offK, int64_t MN,
                                                     int64_t K, bool kcontig, bool sumabs,
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
      m = fmax(m, sumabs ? fabs(x.x) + fabs(x.y) : fmax(fabs(x.x), fabs(x.y)));
    }
  }
  m = wave_max(m);
  if (lane == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

// scale[0] = 2^(15 - e) for max = f 2^e, f in [0.5, 1): the scaled maximum
// lies in [2^14, 2^15); an all-zero operand gets scale 1
__global__ __launch_bounds__(256) void scale_kernel(const double *partials, int nblocks,
                                                    double *scale) {
  __shared__ double sh[256];
  double m = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += 256) m = fmax(m, partials[b]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double mx = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
    int e = 0;
    (void)frexp(mx, &e);
    scale[0] = mx > 0.0 ? ldexp(1.0, 15 - e) : 1.0;
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
  uint4 *out;            // [Kc/2][MNp] pairs of k
  uint2 *out_s;          // or null: [Kc/2][MNp] pairs of k of the words of re + im
  const double *scale;   // FP16: the operand's power-of-two scale
};

// packs a 64(k) x 64(mn) tile: reads in the source's contiguous order,
// transposes through LDS, writes pairs of k mn-fastest; SUMS also writes the
// words of re + im (half parts only)
template <bool FP16, bool SUMS>
__global__ __launch_bounds__(256) void pack_kernel(PackArgs p) {
  static_assert(FP16 || !SUMS, "the re + im plane is packed in half parts");
  __shared__ uint2 tile[kPackTile][kPackTile + 1];
  __shared__ unsigned int tile_s[SUMS ? kPackTile : 1][kPackTile + 1];
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
    unsigned int vs = 0u;
    const int64_t ok = oK[kk], om = oMN[mm];
    if (ok >= 0 && om >= 0) {
      const double2 x = p.X[om + ok];
      if constexpr (FP16) {
        v.x = split_part_f16(sc * x.x);
        v.y = split_part_f16(sc * p.sign * x.y);
        if constexpr (SUMS) vs = split_part_f16(sc * (x.x + p.sign * x.y));
      } else {
        v.x = split_part(x.x);
        v.y = split_part(p.sign * x.y);
      }
    }
    tile[kk][mm] = v;
    if constexpr (SUMS) tile_s[kk][mm] = vs;
  }
  __syncthreads();
  // kb and Kc are even: a pair never straddles the chunk's end
  for (int e = t; e < kPackTile * kPackTile / 2; e += 256) {
    const int mm = e % kPackTile, kp = e / kPackTile;
    const int64_t krow = kb + 2 * kp, mn = mnb + mm;
    if (krow < p.Kc && mn < p.MNp) {
      const uint2 w0 = tile[2 * kp][mm], w1 = tile[2 * kp + 1][mm];
      p.out[(krow >> 1) * p.MNp + mn] = make_uint4(w0.x, w0.y, w1.x, w1.y);
      if constexpr (SUMS)
        p.out_s[(krow >> 1) * p.MNp + mn] = make_uint2(tile_s[2 * kp][mm], tile_s[2 * kp + 1][mm]);
    }
  }
}

struct BgArgs {
  const uint4 *A;  // [Kc/2][Mp] pairs of k
  const uint4 *B;  // [Kc/2][Np] pairs of k
  const uint2 *As, *Bs;  // nine products: the re + im words, [Kc/2][Mp], [Kc/2][Np]
  double *C;
  // split-K partials [splits][M][N]: the raw FP32 sums (float2) of
  // bgemm_kernel, whose scale the reduction applies; FP64 (double2, scaled)
  // from bgemm9_kernel
  void *W;
  const int64_t *rowC, *colC;
  int64_t M, N, Mp, Np, Kc, kChunk;
  double alpha, beta;
  int tilesM, tilesN;
  const double *scaleA, *scaleB;  // FP16: the operands' scales
};

// --- 16 products per complex multiply-add: one complex k per lane ---------
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

// --- 12 products per complex multiply-add: four complex k per half-wave ---
// 32-bit words hold two 16-bit k-slots, the lower half first.  Per complex
// k (word (x, y) = (rh | rl<<16, ih | il<<16)) three words of each operand:
//   A:    (ah, ah)  (al, aih)   (aih, ail)
//   B_re: (bh, bl)  (bh, -bih)  (-bil, -bih)
//   B_im: (bih, bil) (bih, bh)  (bl, bh)
// Twelve words (four k) are the k-slots of three MFMAs (four words each).
struct Ops12 {
  u32x4 v[3];
};
__device__ __forceinline__ unsigned int lo_lo(unsigned int x) {
  return __builtin_amdgcn_perm(x, x, 0x01000100u);
}
// (x >> 16) | (y << 16)
__device__ __forceinline__ unsigned int hi_lo(unsigned int y, unsigned int x) {
  return __builtin_amdgcn_alignbit(y, x, 16);
}
// (x & 0xffff) | (y << 16)
__device__ __forceinline__ unsigned int lo_pair(unsigned int y, unsigned int x) {
  return __builtin_amdgcn_perm(y, x, 0x05040100u);
}
__device__ __forceinline__ void put12(Ops12 &o, int w, unsigned int v) {
  o.v[w >> 2][w & 3] = v;
}
// the four k of two pairs p = (k0, k1), q = (k2, k3)
__device__ __forceinline__ Ops12 expand12_a(uint4 p, uint4 q) {
  Ops12 o;
  const unsigned int x[4] = {p.x, p.z, q.x, q.z}, y[4] = {p.y, p.w, q.y, q.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    put12(o, 3 * k, lo_lo(x[k]));
    put12(o, 3 * k + 1, hi_lo(y[k], x[k]));
    put12(o, 3 * k + 2, y[k]);
  }
  return o;
}
__device__ __forceinline__ void expand12_b(uint4 p, uint4 q, Ops12 &re, Ops12 &im) {
  const unsigned int x[4] = {p.x, p.z, q.x, q.z}, y[4] = {p.y, p.w, q.y, q.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    put12(re, 3 * k, x[k]);
    put12(re, 3 * k + 1, lo_pair(y[k], x[k]) ^ 0x80000000u);
    put12(re, 3 * k + 2, hi_lo(y[k], y[k]) ^ 0x80008000u);
    put12(im, 3 * k, y[k]);
    put12(im, 3 * k + 1, lo_pair(x[k], y[k]));
    put12(im, 3 * k + 2, hi_lo(x[k], x[k]));
  }
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

// NT/64 waves over a BM x BN complex tile; each wave owns WM x WN =
// (WM/32) x (WN/32) blocks of 32x32, re and im accumulators.  LDS holds two
// K-stages of pairs of k (16 bytes per element pair).  MINW is the launch
// bound's minimum of waves per SIMD: 2 with 256 threads, so that two
// workgroups share a CU and one's LDS stores and barriers overlap the
// other's multiplies.
template <int NT, int BM, int BN, int BK, int WM, int WN, bool FP16, int TERMS, int MINW>
__global__ __launch_bounds__(NT, MINW) void bgemm_kernel(BgArgs p) {
  constexpr int MB = WM / 32, NB = WN / 32;
  constexpr int WAVES_N = BN / WN;
  constexpr int BP = BK / 2;                     // pairs of k per stage
  static_assert((BM / WM) * (BN / WN) * 64 == NT, "one wave per WM x WN block");
  static_assert(BK % 8 == 0, "eight complex k per 12-product step");
  static_assert(TERMS == 12 || TERMS == 16, "12 or 16 products");
  constexpr int QA = BP * BM / NT;  // pairs per thread and stage
  constexpr int QB = BP * BN / NT;
  static_assert(QA >= 1 && QB >= 1 && NT % BM == 0 && NT % BN == 0, "staging map");

  __shared__ uint4 As[2][BP][BM];
  __shared__ uint4 Bs[2][BP][BN];

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

  // staging registers, named so that the compiler does not promote them to
  // an LDS array (it does so with private arrays while LDS is free)
  static_assert(QA <= 4 && QB <= 4, "staging is written out for up to 4 + 4 pairs");
  uint4 sa0, sa1, sa2, sa3, sb0, sb1, sb2, sb3;
  constexpr int AR = NT / BM, BR = NT / BN;  // pair rows per pass
  const int ap = t / BM, am = t % BM;
  const int bp = t / BN, bn = t % BN;
  auto load = [&](int s) {
    const int64_t pb = (kbeg + static_cast<int64_t>(s) * BK) >> 1;
    const uint4 *a = p.A + (pb + ap) * p.Mp + m0 + am;
    const uint4 *b = p.B + (pb + bp) * p.Np + n0 + bn;
    sa0 = a[0];
    if constexpr (QA > 1) sa1 = a[AR * p.Mp];
    if constexpr (QA > 2) sa2 = a[2 * AR * p.Mp];
    if constexpr (QA > 3) sa3 = a[3 * AR * p.Mp];
    sb0 = b[0];
    if constexpr (QB > 1) sb1 = b[BR * p.Np];
    if constexpr (QB > 2) sb2 = b[2 * BR * p.Np];
    if constexpr (QB > 3) sb3 = b[3 * BR * p.Np];
  };
  auto store = [&](int buf) {
    As[buf][ap][am] = sa0;
    if constexpr (QA > 1) As[buf][ap + AR][am] = sa1;
    if constexpr (QA > 2) As[buf][ap + 2 * AR][am] = sa2;
    if constexpr (QA > 3) As[buf][ap + 3 * AR][am] = sa3;
    Bs[buf][bp][bn] = sb0;
    if constexpr (QB > 1) Bs[buf][bp + BR][bn] = sb1;
    if constexpr (QB > 2) Bs[buf][bp + 2 * BR][bn] = sb2;
    if constexpr (QB > 3) Bs[buf][bp + 3 * BR][bn] = sb3;
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
  const int ra = wm * WM + (lane & 31), rb = wn * WN + (lane & 31);
  for (int s = 0; s < nstages; ++s) {
    const int buf = s & 1;
    // stage s+1 goes from registers to the other LDS buffer (read in stage
    // s-1, released by that stage's barrier) and stage s+2 is loaded, half
    // way through this stage's MFMAs: the wave keeps issuing MFMAs around
    // its memory instructions instead of stopping for them before the
    // barrier (8.6 % less time per K-stage, measured)
    auto advance = [&]() {
      if (s + 1 < nstages) {
        store(buf ^ 1);
        if (s + 2 < nstages) load(s + 2);
      }
    };
    static_assert(MB == 2 && NB == 2, "fragments are written out for 2 x 2 blocks");
    if constexpr (TERMS == 12) {
#pragma unroll
      for (int q = 0; q < BK / 8; ++q) {
        if (q == BK / 16) advance();
        // lanes 0-31: k = 8q .. 8q+3, lanes 32-63: k = 8q+4 .. 8q+7
        const int pl = 4 * q + 2 * (lane >> 5);
        const Ops12 a0 = expand12_a(As[buf][pl][ra], As[buf][pl + 1][ra]);
        const Ops12 a1 = expand12_a(As[buf][pl][ra + 32], As[buf][pl + 1][ra + 32]);
        Ops12 br0, bi0, br1, bi1;
        expand12_b(Bs[buf][pl][rb], Bs[buf][pl + 1][rb], br0, bi0);
        expand12_b(Bs[buf][pl][rb + 32], Bs[buf][pl + 1][rb + 32], br1, bi1);
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          accR[0][0] = mfma_split<FP16>(a0.v[m], br0.v[m], accR[0][0]);
          accI[0][0] = mfma_split<FP16>(a0.v[m], bi0.v[m], accI[0][0]);
          accR[0][1] = mfma_split<FP16>(a0.v[m], br1.v[m], accR[0][1]);
          accI[0][1] = mfma_split<FP16>(a0.v[m], bi1.v[m], accI[0][1]);
          accR[1][0] = mfma_split<FP16>(a1.v[m], br0.v[m], accR[1][0]);
          accI[1][0] = mfma_split<FP16>(a1.v[m], bi0.v[m], accI[1][0]);
          accR[1][1] = mfma_split<FP16>(a1.v[m], br1.v[m], accR[1][1]);
          accI[1][1] = mfma_split<FP16>(a1.v[m], bi1.v[m], accI[1][1]);
        }
      }
    } else {
#pragma unroll
      for (int kp = 0; kp < BP; ++kp) {
        if (kp == BP / 2) advance();
        // lanes 0-31: k = 2kp (first word of the pair), lanes 32-63: 2kp+1
        const int h = lane >> 5;
        const uint2 *Ap = reinterpret_cast<const uint2 *>(&As[buf][kp][0]);
        const uint2 *Bp = reinterpret_cast<const uint2 *>(&Bs[buf][kp][0]);
        // named registers (not arrays), so no fragment is ever placed in LDS
        const u32x4 a0 = expand_a(Ap[2 * ra + h]);
        const u32x4 a1 = expand_a(Ap[2 * (ra + 32) + h]);
        const uint2 c0 = Bp[2 * rb + h];
        const uint2 c1 = Bp[2 * (rb + 32) + h];
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
        if (p.W) {
          // an FP32 sum is stored exactly as it is; the power-of-two unscale
          // is applied to the FP64 sum of the splits
          const int64_t o = (static_cast<int64_t>(blockIdx.y) * p.M + m) * p.N + n;
          reinterpret_cast<float2 *>(p.W)[o] = make_float2(accR[i][j][r], accI[i][j][r]);
        } else {
          const double vr = unscale * static_cast<double>(accR[i][j][r]);
          const double vi = unscale * static_cast<double>(accI[i][j][r]);
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

// C = alpha / (scaleA scaleB) * sum_s W[s] + beta * C for the FP32 split-K
// partials of bgemm_kernel, summed in FP64 in split order (the same result,
// bit for bit, as summing the scaled partials: the scales are powers of two).
// Threads run along n (coalesced partials), blocks stride over m.
__global__ __launch_bounds__(256) void split_reduce_kernel(
    const float2 *W, double2 *C, const int64_t *rowC, const int64_t *colC, int64_t M,
    int64_t N, int splits, double alpha, double beta, const double *scaleA,
    const double *scaleB) {
  const int64_t n = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (n >= N) return;
  const double f = scaleA ? alpha * (1.0 / (scaleA[0] * scaleB[0])) : alpha;
  const int64_t total = M * N;
  const int64_t cn = colC[n];
  for (int64_t m = blockIdx.y; m < M; m += gridDim.y) {
    const int64_t e = m * N + n;
    double sr = 0.0, si = 0.0;
    for (int s = 0; s < splits; ++s) {
      const float2 v = W[s * total + e];
      sr += static_cast<double>(v.x);
      si += static_cast<double>(v.y);
    }
    double2 *c = C + rowC[m] + cn;
    double2 out = make_double2(f * sr, f * si);
    if (beta != 0.0) {
      const double2 old = *c;
      out.x += beta * old.x;
      out.y += beta * old.y;
    }
    *c = out;
  }
}

// --- 9 products per complex multiply-add (three-product complex form) -----
//   P1 = sum ar br, P2 = sum ai bi, P3 = sum (ar + ai)(br + bi),
//   Re = P1 - P2, Im = P3 - P1 - P2 (formed in FP64 from the FP32 sums),
// each real product split as hi*hi + hi*lo + lo*hi: 9 half products per
// complex MAC instead of 12.  s = ar + ai is split from its FP64 value (the
// operand's scale keeps |ar| + |ai| inside half's range).  Per K-stage (16
// complex k) a half-wave takes 8 k; one part's 8 words (x_k = h | l << 16)
// become 12 operand words, 24 k-slots = 3 MFMAs:
//   A: [ah ah al] per k:  (ah0 ah0) (al0 ah1) (ah1 al1)  per two k
//   B: [bh bl bh] per k:  (bh0 bl0) (bh0 bh1) (bl1 bh1)  per two k
struct Ops9 {
  u32x4 v[3];
};
__device__ __forceinline__ Ops9 expand9_a(const unsigned int (&x)[8]) {
  Ops9 o;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    put12(*reinterpret_cast<Ops12 *>(&o), 3 * e, lo_lo(x[2 * e]));
    put12(*reinterpret_cast<Ops12 *>(&o), 3 * e + 1, hi_lo(x[2 * e + 1], x[2 * e]));
    put12(*reinterpret_cast<Ops12 *>(&o), 3 * e + 2, x[2 * e + 1]);
  }
  return o;
}
__device__ __forceinline__ Ops9 expand9_b(const unsigned int (&y)[8]) {
  Ops9 o;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    put12(*reinterpret_cast<Ops12 *>(&o), 3 * e, y[2 * e]);
    put12(*reinterpret_cast<Ops12 *>(&o), 3 * e + 1, lo_pair(y[2 * e + 1], y[2 * e]));
    put12(*reinterpret_cast<Ops12 *>(&o), 3 * e + 2, hi_lo(y[2 * e + 1], y[2 * e + 1]));
  }
  return o;
}

// 4 waves of 64 x 64 on a 128 x 128 tile, one workgroup per CU (three
// accumulator sets and the raw words of two k-blocks fill the register file
// of one wave per SIMD); LDS holds two K-stages of the (re, im) pairs and the
// re + im pairs.
template <int NT, int BM, int BN, int BK, int WM, int WN>
__global__ __launch_bounds__(NT, 1) void bgemm9_kernel(BgArgs p) {
  constexpr int MB = WM / 32, NB = WN / 32;
  constexpr int WAVES_N = BN / WN;
  constexpr int BP = BK / 2;
  static_assert((BM / WM) * (BN / WN) * 64 == NT, "one wave per WM x WN block");
  static_assert(BK == 16 && MB == 2 && NB == 2, "one 16-k step per stage, 2 x 2 blocks");
  constexpr int QA = BP * BM / NT, QB = BP * BN / NT;
  static_assert(QA == 4 && QB == 4 && NT % BM == 0 && NT % BN == 0, "staging map");

  __shared__ uint4 As[2][BP][BM];
  __shared__ uint4 Bs[2][BP][BN];
  __shared__ uint2 Av[2][BP][BM];
  __shared__ uint2 Bv[2][BP][BN];

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

  // staging registers, named (a private array would be promoted to LDS)
  uint4 sa0, sa1, sa2, sa3, sb0, sb1, sb2, sb3;
  uint2 va0, va1, va2, va3, vb0, vb1, vb2, vb3;
  constexpr int AR = NT / BM, BR = NT / BN;
  const int ap = t / BM, am = t % BM;
  const int bp = t / BN, bn = t % BN;
  auto load = [&](int s) {
    const int64_t pb = (kbeg + static_cast<int64_t>(s) * BK) >> 1;
    const int64_t oa = (pb + ap) * p.Mp + m0 + am, ob = (pb + bp) * p.Np + n0 + bn;
    const int64_t sa = AR * p.Mp, sbs = BR * p.Np;
    sa0 = p.A[oa]; sa1 = p.A[oa + sa]; sa2 = p.A[oa + 2 * sa]; sa3 = p.A[oa + 3 * sa];
    va0 = p.As[oa]; va1 = p.As[oa + sa]; va2 = p.As[oa + 2 * sa]; va3 = p.As[oa + 3 * sa];
    sb0 = p.B[ob]; sb1 = p.B[ob + sbs]; sb2 = p.B[ob + 2 * sbs]; sb3 = p.B[ob + 3 * sbs];
    vb0 = p.Bs[ob]; vb1 = p.Bs[ob + sbs]; vb2 = p.Bs[ob + 2 * sbs]; vb3 = p.Bs[ob + 3 * sbs];
  };
  auto store = [&](int buf) {
    As[buf][ap][am] = sa0; As[buf][ap + AR][am] = sa1;
    As[buf][ap + 2 * AR][am] = sa2; As[buf][ap + 3 * AR][am] = sa3;
    Av[buf][ap][am] = va0; Av[buf][ap + AR][am] = va1;
    Av[buf][ap + 2 * AR][am] = va2; Av[buf][ap + 3 * AR][am] = va3;
    Bs[buf][bp][bn] = sb0; Bs[buf][bp + BR][bn] = sb1;
    Bs[buf][bp + 2 * BR][bn] = sb2; Bs[buf][bp + 3 * BR][bn] = sb3;
    Bv[buf][bp][bn] = vb0; Bv[buf][bp + BR][bn] = vb1;
    Bv[buf][bp + 2 * BR][bn] = vb2; Bv[buf][bp + 3 * BR][bn] = vb3;
  };

  const int lane = t & 63;
  const int w = t >> 6;
  const int wm = w / WAVES_N, wn = w % WAVES_N;
  // P1, P2, P3 of each 32x32 block
  f32x16 acc1[MB][NB], acc2[MB][NB], acc3[MB][NB];
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        acc1[i][j][r] = 0.f;
        acc2[i][j][r] = 0.f;
        acc3[i][j][r] = 0.f;
      }

  if (nstages > 0) {
    load(0);
    store(0);
    if (nstages > 1) load(1);
    __syncthreads();
  }
  const int ra = wm * WM + (lane & 31), rb = wn * WN + (lane & 31);
  const int pl = 4 * (lane >> 5);  // lanes 0-31: k 0-7 of the stage, 32-63: k 8-15
  for (int s = 0; s < nstages; ++s) {
    const int buf = s & 1;
    // the stage's words of this half-wave (8 k) for both M-blocks and both
    // N-blocks, read from LDS once: xa[part][block] / yb[part][block], part
    // 0 re, 1 im, 2 re + im
    // (the re + im words last, so that the first products need not wait for them)
    unsigned int xa[3][MB][8], yb[3][NB][8];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
#pragma unroll
      for (int i = 0; i < MB; ++i) {
        const uint4 q = As[buf][pl + e][ra + 32 * i];
        xa[0][i][2 * e] = q.x; xa[0][i][2 * e + 1] = q.z;
        xa[1][i][2 * e] = q.y; xa[1][i][2 * e + 1] = q.w;
      }
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const uint4 q = Bs[buf][pl + e][rb + 32 * j];
        yb[0][j][2 * e] = q.x; yb[0][j][2 * e + 1] = q.z;
        yb[1][j][2 * e] = q.y; yb[1][j][2 * e + 1] = q.w;
      }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
#pragma unroll
      for (int i = 0; i < MB; ++i) {
        const uint2 v = Av[buf][pl + e][ra + 32 * i];
        xa[2][i][2 * e] = v.x; xa[2][i][2 * e + 1] = v.y;
      }
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const uint2 v = Bv[buf][pl + e][rb + 32 * j];
        yb[2][j][2 * e] = v.x; yb[2][j][2 * e + 1] = v.y;
      }
    }
    // stage s+1 goes to the other LDS buffer (read in stage s-1, released by
    // its barrier) and stage s+2 is loaded, before the products
    if (s + 1 < nstages) {
      store(buf ^ 1);
      if (s + 2 < nstages) load(s + 2);
    }
    auto product = [&](int part, f32x16 (&acc)[MB][NB]) {
      const Ops9 a0 = expand9_a(xa[part][0]), a1 = expand9_a(xa[part][1]);
      const Ops9 b0 = expand9_b(yb[part][0]), b1 = expand9_b(yb[part][1]);
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        acc[0][0] = mfma_split<true>(a0.v[m], b0.v[m], acc[0][0]);
        acc[0][1] = mfma_split<true>(a0.v[m], b1.v[m], acc[0][1]);
        acc[1][0] = mfma_split<true>(a1.v[m], b0.v[m], acc[1][0]);
        acc[1][1] = mfma_split<true>(a1.v[m], b1.v[m], acc[1][1]);
      }
    };
    product(0, acc1);
    product(1, acc2);
    product(2, acc3);
    __syncthreads();
  }

  // lane l, register r of a 32x32 block: row 8*(r/4) + 4*(l/32) + r%4, column l%32
  const double unscale = 1.0 / (p.scaleA[0] * p.scaleB[0]);
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
        const double p1 = static_cast<double>(acc1[i][j][r]);
        const double p2 = static_cast<double>(acc2[i][j][r]);
        const double p3 = static_cast<double>(acc3[i][j][r]);
        const double vr = unscale * (p1 - p2);
        const double vi = unscale * (p3 - (p1 + p2));
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

// multiply tiles: 1 = 8 waves on 128 x 256 (one workgroup per CU, 96 KB of
// LDS); 2 (default) = 4 waves on 128 x 128 (two workgroups per CU, 64 KB
// each: one's LDS stores and barriers overlap the other's MFMAs; 3.3 %
// faster over a No=64 Nv=512 split pass); both with 64 x 64 wave tiles
constexpr int kXBK = 16, kXWM = 64, kXWN = 64;
int g_mixed_tile = 2;

template <int NT, int BM, int BN, int MINW>
void dispatch_bgemm(const BgArgs &a, dim3 grid, bool half, int terms, hipStream_t s) {
  if (half && terms == 12)
    bgemm_kernel<NT, BM, BN, kXBK, kXWM, kXWN, true, 12, MINW><<<grid, NT, 0, s>>>(a);
  else if (half)
    bgemm_kernel<NT, BM, BN, kXBK, kXWM, kXWN, true, 16, MINW><<<grid, NT, 0, s>>>(a);
  else if (terms == 12)
    bgemm_kernel<NT, BM, BN, kXBK, kXWM, kXWN, false, 12, MINW><<<grid, NT, 0, s>>>(a);
  else
    bgemm_kernel<NT, BM, BN, kXBK, kXWM, kXWN, false, 16, MINW><<<grid, NT, 0, s>>>(a);
}

}  // namespace

void set_mixed_tile(int t) { g_mixed_tile = t == 1 ? 1 : 2; }

MixedTile mixed_tile(int terms) {
  return g_mixed_tile == 2 || terms == 9 ? MixedTile{128, 128, kXBK}
                                         : MixedTile{128, 256, kXBK};
}

void launch_split_scale(const double *X, const int64_t *offMN, const int64_t *offK,
                        int64_t MN, int64_t K, bool kcontig, bool sumabs,
                        double *partials, double *scale, hipStream_t s) {
  const int64_t n = MN * K;
  const int blocks = static_cast<int>(std::max<int64_t>(
      1, std::min<int64_t>(kMaxBlocks, (n + 255) / 256)));
  maxabs_kernel<<<blocks, 256, 0, s>>>(reinterpret_cast<const double2 *>(X), offMN, offK,
                                       MN, K, kcontig, sumabs, partials);
  CC4S_HIP(hipGetLastError());
  scale_kernel<<<1, 256, 0, s>>>(partials, blocks, scale);
  CC4S_HIP(hipGetLastError());
}

size_t split_scale_partials_bytes() { return kMaxBlocks * sizeof(double); }

void launch_split_reduce(const void *W, double *C, const int64_t *rowC,
                         const int64_t *colC, int64_t M, int64_t N, int splits,
                         double alpha, double beta, const double *scaleA,
                         const double *scaleB, hipStream_t s) {
  if (M <= 0 || N <= 0) return;
  const dim3 grid(static_cast<unsigned>((N + 255) / 256),
                  static_cast<unsigned>(std::min<int64_t>(M, 4096)));
  split_reduce_kernel<<<grid, 256, 0, s>>>(static_cast<const float2 *>(W),
                                           reinterpret_cast<double2 *>(C), rowC, colC, M,
                                           N, splits, alpha, beta, scaleA, scaleB);
  CC4S_HIP(hipGetLastError());
}

void launch_pack(const double *X, const int64_t *offMN, const int64_t *offK,
                 int64_t MN, int64_t K, int64_t k0, int64_t Kc, int64_t MNp,
                 bool conj, bool kcontig, void *out, void *out_s,
                 const double *scale, hipStream_t s) {
  CC4S_ASSERT(out_s == nullptr || scale != nullptr, "the re + im plane is half only");
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
  a.out = static_cast<uint4 *>(out);
  a.out_s = static_cast<uint2 *>(out_s);
  a.scale = scale;
  dim3 grid(static_cast<unsigned>((MNp + kPackTile - 1) / kPackTile),
            static_cast<unsigned>((Kc + kPackTile - 1) / kPackTile));
  if (out_s)
    pack_kernel<true, true><<<grid, 256, 0, s>>>(a);
  else if (scale)
    pack_kernel<true, false><<<grid, 256, 0, s>>>(a);
  else
    pack_kernel<false, false><<<grid, 256, 0, s>>>(a);
  CC4S_HIP(hipGetLastError());
}

void launch_bgemm(const void *Apack, const void *Bpack, double *C, void *W,
                  const int64_t *rowC, const int64_t *colC, int64_t M, int64_t N,
                  int64_t Mp, int64_t Np, int64_t Kc, int splits, double alpha,
                  double beta, const double *scaleA, const double *scaleB,
                  int terms, const void *Asum, const void *Bsum, hipStream_t s) {
  CC4S_ASSERT((scaleA == nullptr) == (scaleB == nullptr),
              "split-half operands need both scales");
  CC4S_ASSERT(terms == 9 || terms == 12 || terms == 16,
              "split products per complex MAC: 9, 12 or 16");
  CC4S_ASSERT(terms != 9 || (scaleA && Asum && Bsum),
              "9 products need half parts and the re + im planes");
  const MixedTile T = mixed_tile(terms);
  CC4S_ASSERT(Kc % T.BK == 0 && Mp % T.BM == 0 && Np % T.BN == 0,
              "packed operands are padded to whole tiles");
  BgArgs a;
  a.scaleA = scaleA;
  a.scaleB = scaleB;
  a.A = static_cast<const uint4 *>(Apack);
  a.B = static_cast<const uint4 *>(Bpack);
  a.As = static_cast<const uint2 *>(Asum);
  a.Bs = static_cast<const uint2 *>(Bsum);
  a.C = C;
  a.W = splits > 1 ? W : nullptr;
  a.rowC = rowC;
  a.colC = colC;
  a.M = M; a.N = N; a.Mp = Mp; a.Np = Np; a.Kc = Kc;
  const int64_t stages = Kc / kXBK;
  a.kChunk = ((stages + splits - 1) / splits) * kXBK;
  a.alpha = alpha;
  a.beta = beta;
  a.tilesM = static_cast<int>(Mp / T.BM);
  a.tilesN = static_cast<int>(Np / T.BN);
  dim3 grid(a.tilesM * a.tilesN, splits);
  if (terms == 9)
    bgemm9_kernel<256, 128, 128, kXBK, 64, 64><<<grid, 256, 0, s>>>(a);
  else if (g_mixed_tile == 2)
    dispatch_bgemm<256, 128, 128, 2>(a, grid, scaleA != nullptr, terms, s);
  else
    dispatch_bgemm<512, 128, 256, 1>(a, grid, scaleA != nullptr, terms, s);
  CC4S_HIP(hipGetLastError());
}

}  // namespace cc4s_amd
