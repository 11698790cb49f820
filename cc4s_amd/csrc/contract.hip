// GETT kernels (FP64 MFMA, gfx950) and the contraction planner.
#include "contract.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <sstream>
#include <type_traits>

namespace cc4s_amd {

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;  // 4 wave64 per workgroup
constexpr int kPad = 16;       // LDS row padding (doubles)

struct KArgs {
  const double *A;
  const double *B;
  double *C;
  double *W;
  const int64_t *rowA, *colA, *rowB, *colB, *rowC, *colC;
  int64_t M, N, K, kChunk;
  double alpha, beta;
  double signA, signB;  // -1 conjugates the operand's imaginary part
  int tilesM, tilesN;
  // second K segment (SEG2 kernels): k in [K1p, K1p + K2) reads
  // A2[rowA2[m] + colA2[k - K1p]] and B2[rowB2[k - K1p] + colB2[n]], times
  // f2, with its own conjugation signs; the first segment is k < K1 (K1p is
  // K1 rounded up to whole K-stages)
  const double *A2, *B2;
  const int64_t *rowA2, *colA2, *rowB2, *colB2;
  int64_t K1, K1p, K2;
  double f2, signA2, signB2;
};

// Workgroup tile BM x BN, K staged BK at a time through LDS; 4 waves in a
// (BM/WM) x (BN/WN) grid, each owning WM x WN of C as 16x16 MFMA blocks.
// A is read either along K (AKC: consecutive lanes -> consecutive k) or along
// M; B along K (BKC) or along N.  LDS holds A as [k][m] and B as [k][n] (one
// plane per real/imaginary part), which is exactly the lane order of the
// MFMA operands (lane l reads row l%16 at k = l/16).
template <bool CPLX, int BM, int BN, int BK, int WM, int WN, bool AKC, bool BKC>
__global__ __launch_bounds__(kThreads, 2) void gett_kernel(KArgs p) {
  constexpr int NPL = CPLX ? 2 : 1;
  constexpr int MB = WM / 16, NB = WN / 16;
  constexpr int WAVES_N = BN / WN;
  static_assert((BM / WM) * (BN / WN) == 4, "four waves per workgroup");
  static_assert(BK % 4 == 0, "BK multiple of the MFMA k");
  constexpr int EA = BM * BK / kThreads;
  constexpr int EB = BN * BK / kThreads;
  static_assert(EA >= 1 && EB >= 1, "tile too small");
  static_assert(AKC ? (kThreads % BK == 0) : (kThreads % BM == 0), "A map");
  static_assert(BKC ? (kThreads % BK == 0) : (kThreads % BN == 0), "B map");
  constexpr int NKA = AKC ? 1 : EA;  // distinct k per thread per stage
  constexpr int NRA = AKC ? EA : 1;  // distinct m per thread
  constexpr int NKB = BKC ? 1 : EB;
  constexpr int NRB = BKC ? EB : 1;

  __shared__ double As[NPL][BK][BM + kPad];
  __shared__ double Bs[NPL][BK][BN + kPad];

  const int t = threadIdx.x;
  // tile index: spread consecutive workgroups (round-robin over the 8 XCDs)
  // back into neighbouring tiles per XCD, then walk groups of 8 M-tiles so
  // that concurrently running tiles share A rows and B columns in L2.
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
  const int64_t kend = min(p.K, kbeg + p.kChunk);
  const int nstages = kend > kbeg ? static_cast<int>((kend - kbeg + BK - 1) / BK) : 0;

  // fixed per-thread rows of A and columns of B
  int64_t aRow[NRA];
  bool aRowOk[NRA];
#pragma unroll
  for (int i = 0; i < NRA; ++i) {
    const int64_t m = m0 + (AKC ? (t / BK + i * (kThreads / BK)) : (t % BM));
    aRowOk[i] = m < p.M;
    aRow[i] = aRowOk[i] ? p.rowA[m] : 0;
  }
  int64_t bCol[NRB];
  bool bColOk[NRB];
#pragma unroll
  for (int i = 0; i < NRB; ++i) {
    const int64_t n = n0 + (BKC ? (t / BK + i * (kThreads / BK)) : (t % BN));
    bColOk[i] = n < p.N;
    bCol[i] = bColOk[i] ? p.colB[n] : 0;
  }
  auto a_k = [&](int i) { return AKC ? (t % BK) : (t / BM + i * (kThreads / BM)); };
  auto b_k = [&](int i) { return BKC ? (t % BK) : (t / BN + i * (kThreads / BN)); };

  // K offsets are prefetched one stage ahead of the data they address
  int64_t aK[NKA], aKn[NKA], bK[NKB], bKn[NKB];
  bool aKok[NKA], aKokn[NKA], bKok[NKB], bKokn[NKB];
  auto load_koff = [&](int s, int64_t *ak, bool *akok, int64_t *bk, bool *bkok) {
    const int64_t kb = kbeg + static_cast<int64_t>(s) * BK;
#pragma unroll
    for (int i = 0; i < NKA; ++i) {
      const int64_t k = kb + a_k(i);
      akok[i] = k < kend;
      ak[i] = akok[i] ? p.colA[k] : 0;
    }
#pragma unroll
    for (int i = 0; i < NKB; ++i) {
      const int64_t k = kb + b_k(i);
      bkok[i] = k < kend;
      bk[i] = bkok[i] ? p.rowB[k] : 0;
    }
  };

  double ra[EA][NPL], rb[EB][NPL];
  auto load_data = [&]() {
#pragma unroll
    for (int i = 0; i < EA; ++i) {
      const int ir = AKC ? i : 0, ik = AKC ? 0 : i;
      const bool ok = aRowOk[ir] && aKok[ik];
      const int64_t off = aRow[ir] + aK[ik];
      if constexpr (CPLX) {
        double2 v = ok ? *reinterpret_cast<const double2 *>(p.A + 2 * off)
                       : make_double2(0.0, 0.0);
        ra[i][0] = v.x;
        ra[i][1] = v.y * p.signA;
      } else {
        ra[i][0] = ok ? p.A[off] : 0.0;
      }
    }
#pragma unroll
    for (int i = 0; i < EB; ++i) {
      const int ir = BKC ? i : 0, ik = BKC ? 0 : i;
      const bool ok = bColOk[ir] && bKok[ik];
      const int64_t off = bCol[ir] + bK[ik];
      if constexpr (CPLX) {
        double2 v = ok ? *reinterpret_cast<const double2 *>(p.B + 2 * off)
                       : make_double2(0.0, 0.0);
        rb[i][0] = v.x;
        rb[i][1] = v.y * p.signB;
      } else {
        rb[i][0] = ok ? p.B[off] : 0.0;
      }
    }
  };
  auto store_lds = [&]() {
#pragma unroll
    for (int i = 0; i < EA; ++i) {
      const int k = a_k(i);
      const int m = AKC ? (t / BK + i * (kThreads / BK)) : (t % BM);
#pragma unroll
      for (int c = 0; c < NPL; ++c) As[c][k][m] = ra[i][c];
    }
#pragma unroll
    for (int i = 0; i < EB; ++i) {
      const int k = b_k(i);
      const int n = BKC ? (t / BK + i * (kThreads / BK)) : (t % BN);
#pragma unroll
      for (int c = 0; c < NPL; ++c) Bs[c][k][n] = rb[i][c];
    }
  };

  const int lane = t & 63;
  const int w = t >> 6;
  const int wm = w / WAVES_N, wn = w % WAVES_N;
  dbl4 accR[MB][NB], accI[MB][NB];
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      accR[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
      accI[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
    }

  if (nstages > 0) {
    load_koff(0, aK, aKok, bK, bKok);
    if (nstages > 1) load_koff(1, aKn, aKokn, bKn, bKokn);
    load_data();
    store_lds();
    __syncthreads();
  }
  for (int s = 0; s < nstages; ++s) {
    const bool more = s + 1 < nstages;
    if (more) {
#pragma unroll
      for (int i = 0; i < NKA; ++i) { aK[i] = aKn[i]; aKok[i] = aKokn[i]; }
#pragma unroll
      for (int i = 0; i < NKB; ++i) { bK[i] = bKn[i]; bKok[i] = bKokn[i]; }
      if (s + 2 < nstages) load_koff(s + 2, aKn, aKokn, bKn, bKokn);
      load_data();
    }
#pragma unroll
    for (int kq = 0; kq < BK / 4; ++kq) {
      const int kr = kq * 4 + (lane >> 4);
      double ar[MB], ai[MB], br[NB], bi[NB];
#pragma unroll
      for (int i = 0; i < MB; ++i) {
        ar[i] = As[0][kr][wm * WM + i * 16 + (lane & 15)];
        if constexpr (CPLX) ai[i] = As[NPL - 1][kr][wm * WM + i * 16 + (lane & 15)];
      }
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        br[j] = Bs[0][kr][wn * WN + j * 16 + (lane & 15)];
        if constexpr (CPLX) bi[j] = Bs[NPL - 1][kr][wn * WN + j * 16 + (lane & 15)];
      }
#pragma unroll
      for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j)
          accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], accR[i][j], 0, 0, 0);
      if constexpr (CPLX) {
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
          for (int j = 0; j < NB; ++j)
            accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi[j], accI[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < MB; ++i) {
          const double nai = -ai[i];
#pragma unroll
          for (int j = 0; j < NB; ++j)
            accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai, bi[j], accR[i][j], 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
          for (int j = 0; j < NB; ++j)
            accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], br[j], accI[i][j], 0, 0, 0);
      }
    }
    if (more) {
      __syncthreads();
      store_lds();
      __syncthreads();
    }
  }

  // epilogue: lane l, register r of a 16x16 f64 MFMA block holds row
  // 4*r + l/16, column l%16
  int64_t cCol[NB];
  bool cColOk[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const int64_t n = n0 + wn * WN + j * 16 + (lane & 15);
    cColOk[j] = n < p.N;
    cCol[j] = (cColOk[j] && !p.W) ? p.colC[n] : 0;
  }
#pragma unroll
  for (int i = 0; i < MB; ++i) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t m = m0 + wm * WM + i * 16 + 4 * r + (lane >> 4);
      if (m >= p.M) continue;
      const int64_t rc = p.W ? 0 : p.rowC[m];
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        if (!cColOk[j]) continue;
        const int64_t n = n0 + wn * WN + j * 16 + (lane & 15);
        const double vr = accR[i][j][r];
        const double vi = accI[i][j][r];
        if (p.W) {
          const int64_t o = (static_cast<int64_t>(blockIdx.y) * p.M + m) * p.N + n;
          if constexpr (CPLX) {
            reinterpret_cast<double2 *>(p.W)[o] = make_double2(vr, vi);
          } else {
            p.W[o] = vr;
          }
        } else {
          const int64_t o = rc + cCol[j];
          if constexpr (CPLX) {
            double2 *c = reinterpret_cast<double2 *>(p.C) + o;
            double2 out = make_double2(p.alpha * vr, p.alpha * vi);
            if (p.beta != 0.0) {
              const double2 old = *c;
              out.x += p.beta * old.x;
              out.y += p.beta * old.y;
            }
            *c = out;
          } else {
            double *c = p.C + o;
            double out = p.alpha * vr;
            if (p.beta != 0.0) out += p.beta * *c;
            *c = out;
          }
        }
      }
    }
  }
}

// Complex GETT, second generation: NT/64 waves on a BM x BN complex tile,
// one workgroup per CU.  LDS holds two K-stages (double buffer) with (re, im)
// interleaved, so an MFMA operand pair is one 128-bit LDS read and a K-stage
// needs one barrier.  Global loads for stage s+2 are issued while stage s is
// multiplied; their K offsets were fetched a stage earlier still.
//
// M3 selects the three-product complex multiply (Gauss/Karatsuba):
//   P1 = ar br, P2 = ai bi, P3 = (ar + ai)(br + bi),
//   re = P1 - P2, im = P3 - P1 - P2,
// three MFMA products per complex MAC instead of four.  Its error bound is
// the four-product one with |ar|+|ai| and |br|+|bi| in place of |a|, |b|.
constexpr int kZThreads = 512;

// DEEP keeps two stages of global loads in flight (two register sets): the
// loads for stage s+3 are issued while stage s is multiplied.  MINW is the
// launch bound's minimum of waves per SIMD (2 with 256 threads: two
// workgroups share a CU, so one's prologue and epilogue overlap the other's
// multiplies).
//
// SEG2 adds a second K segment with its own operands (two contractions with
// the same result index partition, e.g. a particle and a hole term of (T),
// accumulated in one pass over C); a K-stage lies in one segment.
template <int NT, int BM, int BN, int BK, int WM, int WN, bool AKC, bool BKC, bool M3,
          bool DEEP, int MINW = 1, bool SEG2 = false>
__global__ __launch_bounds__(NT, MINW) void zgett_kernel(KArgs p) {
  constexpr int MB = WM / 16, NB = WN / 16;
  constexpr int WAVES_N = BN / WN;
  static_assert((BM / WM) * (BN / WN) * 64 == NT, "one wave per WM x WN block");
  static_assert(BK % 4 == 0, "BK multiple of the MFMA k");
  constexpr int EA = BM * BK / NT;
  constexpr int EB = BN * BK / NT;
  static_assert(EA >= 1 && EB >= 1, "tile too small");
  static_assert(AKC ? (NT % BK == 0) : (NT % BM == 0), "A map");
  static_assert(BKC ? (NT % BK == 0) : (NT % BN == 0), "B map");
  constexpr int NKA = AKC ? 1 : EA;
  constexpr int NRA = AKC ? EA : 1;
  constexpr int NKB = BKC ? 1 : EB;
  constexpr int NRB = BKC ? EB : 1;
  // complex elements per LDS row: BM+1 shifts consecutive k-rows by one
  // 16-byte bank group, so the column-wise stores of K-contiguous operands
  // and the row-wise 128-bit operand reads are both conflict-free per
  // 16-lane pass
  constexpr int LDA = BM + 1, LDB = BN + 1;

  __shared__ double2 As[2][BK][LDA];
  __shared__ double2 Bs[2][BK][LDB];

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
  const int64_t kend = min(p.K, kbeg + p.kChunk);
  const int nstages = kend > kbeg ? static_cast<int>((kend - kbeg + BK - 1) / BK) : 0;

  auto a_m = [&](int i) { return AKC ? (t / BK + i * (NT / BK)) : (t % BM); };
  auto a_k = [&](int i) { return AKC ? (t % BK) : (t / BM + i * (NT / BM)); };
  auto b_n = [&](int i) { return BKC ? (t / BK + i * (NT / BK)) : (t % BN); };
  auto b_k = [&](int i) { return BKC ? (t % BK) : (t / BN + i * (NT / BN)); };

  int64_t aRow[NRA], aRow2[SEG2 ? NRA : 1];
  bool aRowOk[NRA];
#pragma unroll
  for (int i = 0; i < NRA; ++i) {
    const int64_t m = m0 + a_m(i);
    aRowOk[i] = m < p.M;
    aRow[i] = aRowOk[i] ? p.rowA[m] : 0;
    if constexpr (SEG2) aRow2[i] = aRowOk[i] ? p.rowA2[m] : 0;
  }
  int64_t bCol[NRB], bCol2[SEG2 ? NRB : 1];
  bool bColOk[NRB];
#pragma unroll
  for (int i = 0; i < NRB; ++i) {
    const int64_t n = n0 + b_n(i);
    bColOk[i] = n < p.N;
    bCol[i] = bColOk[i] ? p.colB[n] : 0;
    if constexpr (SEG2) bCol2[i] = bColOk[i] ? p.colB2[n] : 0;
  }

  // K offsets of two stages: stage t's in buffer t & 1, indexed by
  // constants (the loop below is unrolled by two), so no register copies
  // wait for a load
  int64_t aK[2][NKA], bK[2][NKB];
  bool aKok[2][NKA], bKok[2][NKB];
  bool segK[2] = {false, false};  // SEG2: the stage lies in the second segment
  // every load is unconditional (out-of-range lanes read a valid element,
  // kbeg < kend or offset 0, and are zeroed when stored to LDS): a load
  // under a branch makes the compiler wait for it at the branch's end, which
  // would stall the wave a whole K-stage early
  auto load_koff = [&](int s, auto par_c) {
    constexpr int pb = decltype(par_c)::value;
    int64_t *ak = aK[pb], *bk = bK[pb];
    bool *akok = aKok[pb], *bkok = bKok[pb];
    const int64_t kb = kbeg + static_cast<int64_t>(s) * BK;
    if constexpr (SEG2) {
      const bool sg = kb >= p.K1p;
      segK[pb] = sg;
      const int64_t k0 = sg ? p.K1p : 0, klen = sg ? p.K2 : p.K1;
      const int64_t *ca = sg ? p.colA2 : p.colA, *rb = sg ? p.rowB2 : p.rowB;
#pragma unroll
      for (int i = 0; i < NKA; ++i) {
        const int64_t k = kb + a_k(i);
        akok[i] = k < kend && k - k0 < klen;
        ak[i] = ca[akok[i] ? k - k0 : 0];
      }
#pragma unroll
      for (int i = 0; i < NKB; ++i) {
        const int64_t k = kb + b_k(i);
        bkok[i] = k < kend && k - k0 < klen;
        bk[i] = rb[bkok[i] ? k - k0 : 0];
      }
    } else {
#pragma unroll
      for (int i = 0; i < NKA; ++i) {
        const int64_t k = kb + a_k(i);
        akok[i] = k < kend;
        ak[i] = p.colA[akok[i] ? k : kbeg];
      }
#pragma unroll
      for (int i = 0; i < NKB; ++i) {
        const int64_t k = kb + b_k(i);
        bkok[i] = k < kend;
        bk[i] = p.rowB[bkok[i] ? k : kbeg];
      }
    }
  };

  // staging registers; DEEP uses two sets, always indexed by constants
  constexpr int NSET = DEEP ? 2 : 1;
  double2 ra[NSET][EA], rb[NSET][EB];
  bool oka[NSET][EA], okb[NSET][EB];
  bool segS[NSET];  // SEG2: the registers hold a second-segment stage
  auto load_data = [&](auto set_c, auto par_c) {
    constexpr int q = decltype(set_c)::value;
    constexpr int pb = decltype(par_c)::value;
    bool sg = false;
    if constexpr (SEG2) sg = segK[pb];
    segS[q] = sg;
    const double2 *A = reinterpret_cast<const double2 *>(sg ? p.A2 : p.A);
    const double2 *B = reinterpret_cast<const double2 *>(sg ? p.B2 : p.B);
#pragma unroll
    for (int i = 0; i < EA; ++i) {
      const int ir = AKC ? i : 0, ik = AKC ? 0 : i;
      oka[q][i] = aRowOk[ir] && aKok[pb][ik];
      int64_t row = aRow[ir];
      if constexpr (SEG2) row = sg ? aRow2[ir] : row;
      ra[q][i] = A[oka[q][i] ? row + aK[pb][ik] : 0];
    }
#pragma unroll
    for (int i = 0; i < EB; ++i) {
      const int ir = BKC ? i : 0, ik = BKC ? 0 : i;
      okb[q][i] = bColOk[ir] && bKok[pb][ik];
      int64_t col = bCol[ir];
      if constexpr (SEG2) col = sg ? bCol2[ir] : col;
      rb[q][i] = B[okb[q][i] ? col + bK[pb][ik] : 0];
    }
  };
  // out-of-range elements are zeroed and the conjugation signs (and the
  // second segment's factor) applied here, a stage after the loads were
  // issued
  auto store_lds = [&](int buf, auto set_c) {
    constexpr int q = decltype(set_c)::value;
    const double2 zero = make_double2(0.0, 0.0);
    double fa = 1.0, sa = p.signA, sb = p.signB;
    if constexpr (SEG2) {
      fa = segS[q] ? p.f2 : 1.0;
      sa = segS[q] ? p.signA2 : sa;
      sb = segS[q] ? p.signB2 : sb;
    }
#pragma unroll
    for (int i = 0; i < EA; ++i)
      As[buf][a_k(i)][a_m(i)] =
          oka[q][i] ? make_double2(fa * ra[q][i].x, fa * ra[q][i].y * sa) : zero;
#pragma unroll
    for (int i = 0; i < EB; ++i)
      Bs[buf][b_k(i)][b_n(i)] =
          okb[q][i] ? make_double2(rb[q][i].x, rb[q][i].y * sb) : zero;
  };
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, DEEP ? 1 : 0>;
  using P0 = std::integral_constant<int, 0>;
  using P1 = std::integral_constant<int, 1>;

  const int lane = t & 63;
  const int w = t >> 6;
  const int wm = w / WAVES_N, wn = w % WAVES_N;
  // four-product: accR = re, accI = im; M3: accR = P1, accI = P2, accS = P3
  dbl4 accR[MB][NB], accI[MB][NB], accS[M3 ? MB : 1][M3 ? NB : 1];
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      accR[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
      accI[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
      if constexpr (M3) accS[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
    }

  // prologue: stage 0 in LDS buffer 0; stage 1 in registers (DEEP: stage 1
  // in set 1, stage 2 in set 0)
  if (nstages > 0) {
    load_koff(0, P0{});
    if (nstages > 1) load_koff(1, P1{});
    load_data(S0{}, P0{});
    store_lds(0, S0{});
    if (nstages > 1) {
      if (nstages > 2) load_koff(2, P0{});
      load_data(S1{}, P1{});
      if (DEEP && nstages > 2) {
        if (nstages > 3) load_koff(3, P1{});
        load_data(S0{}, P0{});
      }
    }
    __syncthreads();
  }
  // one K-stage: multiply LDS buffer s&1, then move stage s+1 from its
  // registers to the other buffer and load the next stage into them
  auto stage = [&](int s, auto set_c, auto par_c) {
    constexpr int par = decltype(par_c)::value;  // s & 1
    const int buf = s & 1;
    constexpr int AHEAD = DEEP ? 3 : 2;  // stage loaded into the freed registers
    using PD = std::integral_constant<int, (par + AHEAD) & 1>;      // stage s+AHEAD
    using PN = std::integral_constant<int, (par + AHEAD + 1) & 1>;  // stage s+AHEAD+1
    // stage s+1 moves from its registers to the other LDS buffer (read in
    // stage s-1, released by that stage's barrier) and stage s+AHEAD is
    // loaded half way through this stage's MFMAs, so the wave keeps issuing
    // MFMAs around its memory instructions
    auto advance = [&]() {
      if (s + 1 < nstages) {
        store_lds(buf ^ 1, set_c);  // stage s+1
        if (s + AHEAD < nstages) {
          if (s + AHEAD + 1 < nstages) load_koff(s + AHEAD + 1, PN{});
          load_data(set_c, PD{});
        }
      }
    };
#pragma unroll
    for (int kq = 0; kq < BK / 4; ++kq) {
      if (kq == BK / 8) advance();
      const int kr = kq * 4 + (lane >> 4);
      double2 a[MB], b[NB];
#pragma unroll
      for (int i = 0; i < MB; ++i) a[i] = As[buf][kr][wm * WM + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < NB; ++j) b[j] = Bs[buf][kr][wn * WN + j * 16 + (lane & 15)];
      if constexpr (M3) {
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
          for (int j = 0; j < NB; ++j)
            accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].x, b[j].x, accR[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
          for (int j = 0; j < NB; ++j)
            accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].y, b[j].y, accI[i][j], 0, 0, 0);
        double bs[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) bs[j] = b[j].x + b[j].y;
#pragma unroll
        for (int i = 0; i < MB; ++i) {
          const double as = a[i].x + a[i].y;
#pragma unroll
          for (int j = 0; j < NB; ++j)
            accS[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(as, bs[j], accS[i][j], 0, 0, 0);
        }
      } else {
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
          for (int j = 0; j < NB; ++j)
            accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].x, b[j].x, accR[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
          for (int j = 0; j < NB; ++j)
            accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].x, b[j].y, accI[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < MB; ++i) {
          const double nai = -a[i].y;
#pragma unroll
          for (int j = 0; j < NB; ++j)
            accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai, b[j].y, accR[i][j], 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
          for (int j = 0; j < NB; ++j)
            accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].y, b[j].x, accI[i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  };
  // stage s+1 sits in set (s+1)&1 (DEEP) or in the single set
  for (int s = 0; s < nstages; s += 2) {
    stage(s, S1{}, P0{});
    if (s + 1 < nstages) stage(s + 1, S0{}, P1{});
  }

  int64_t cCol[NB];
  bool cColOk[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const int64_t n = n0 + wn * WN + j * 16 + (lane & 15);
    cColOk[j] = n < p.N;
    cCol[j] = (cColOk[j] && !p.W) ? p.colC[n] : 0;
  }
#pragma unroll
  for (int i = 0; i < MB; ++i) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t m = m0 + wm * WM + i * 16 + 4 * r + (lane >> 4);
      if (m >= p.M) continue;
      const int64_t rc = p.W ? 0 : p.rowC[m];
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        if (!cColOk[j]) continue;
        const int64_t n = n0 + wn * WN + j * 16 + (lane & 15);
        double vr = accR[i][j][r];
        double vi = accI[i][j][r];
        if constexpr (M3) {
          const double p1 = vr, p2 = vi;
          vr = p1 - p2;
          vi = accS[i][j][r] - (p1 + p2);
        }
        if (p.W) {
          const int64_t o = (static_cast<int64_t>(blockIdx.y) * p.M + m) * p.N + n;
          reinterpret_cast<double2 *>(p.W)[o] = make_double2(vr, vi);
        } else {
          double2 *c = reinterpret_cast<double2 *>(p.C) + rc + cCol[j];
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

// C[rowC[m] + colC[n]] = alpha * sum_s W[s][m][n] + beta * C, fixed order
template <bool CPLX>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(
    const double *W, double *C, const int64_t *rowC, const int64_t *colC,
    int64_t M, int64_t N, int splits, double alpha, double beta) {
  const int64_t total = M * N;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < total;
       e += static_cast<int64_t>(gridDim.x) * 256) {
    const int64_t m = e / N, n = e - m * N;
    const int64_t o = rowC[m] + colC[n];
    if constexpr (CPLX) {
      const double2 *w = reinterpret_cast<const double2 *>(W);
      double sr = 0.0, si = 0.0;
      for (int s = 0; s < splits; ++s) {
        const double2 v = w[s * total + e];
        sr += v.x;
        si += v.y;
      }
      double2 *c = reinterpret_cast<double2 *>(C) + o;
      double2 out = make_double2(alpha * sr, alpha * si);
      if (beta != 0.0) {
        const double2 old = *c;
        out.x += beta * old.x;
        out.y += beta * old.y;
      }
      *c = out;
    } else {
      double sr = 0.0;
      for (int s = 0; s < splits; ++s) sr += W[s * total + e];
      double out = alpha * sr;
      if (beta != 0.0) out += beta * C[o];
      C[o] = out;
    }
  }
}

constexpr int kMaxRank = 8;
struct SumArgs {
  const double *A;
  double *C;
  int rank;
  int64_t lens[kMaxRank];
  int64_t sA[kMaxRank];
  int64_t sC[kMaxRank];
  int64_t total;
  double alpha, beta, signA;
};

// C = alpha * op(A) + beta * C over C's index space (first index fastest)
template <bool CPLX>
__global__ __launch_bounds__(256) void sum_kernel(SumArgs p) {
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < p.total;
       e += static_cast<int64_t>(gridDim.x) * 256) {
    int64_t r = e, oa = 0, oc = 0;
    for (int d = 0; d < p.rank; ++d) {
      const int64_t q = r / p.lens[d];
      const int64_t i = r - q * p.lens[d];
      r = q;
      oa += i * p.sA[d];
      oc += i * p.sC[d];
    }
    if constexpr (CPLX) {
      double2 a = p.A ? reinterpret_cast<const double2 *>(p.A)[oa]
                      : make_double2(0.0, 0.0);
      double2 *c = reinterpret_cast<double2 *>(p.C) + oc;
      double2 out = make_double2(p.alpha * a.x, p.alpha * p.signA * a.y);
      if (p.beta != 0.0) {
        const double2 old = *c;
        out.x += p.beta * old.x;
        out.y += p.beta * old.y;
      }
      *c = out;
    } else {
      double out = p.A ? p.alpha * p.A[oa] : 0.0;
      if (p.beta != 0.0) out += p.beta * p.C[oc];
      p.C[oc] = out;
    }
  }
}

template <bool CPLX, int BM, int BN, int BK, int WM, int WN>
void dispatch_gett(const GettLaunch &L, const KArgs &args, dim3 grid, hipStream_t s) {
  if (L.a_kcontig && L.b_kcontig)
    gett_kernel<CPLX, BM, BN, BK, WM, WN, true, true><<<grid, kThreads, 0, s>>>(args);
  else if (L.a_kcontig)
    gett_kernel<CPLX, BM, BN, BK, WM, WN, true, false><<<grid, kThreads, 0, s>>>(args);
  else if (L.b_kcontig)
    gett_kernel<CPLX, BM, BN, BK, WM, WN, false, true><<<grid, kThreads, 0, s>>>(args);
  else
    gett_kernel<CPLX, BM, BN, BK, WM, WN, false, false><<<grid, kThreads, 0, s>>>(args);
}

template <int NT, int BM, int BN, int BK, int WM, int WN, bool M3, bool DEEP = false,
          int MINW = 1, bool SEG2 = false>
void dispatch_zgett(const GettLaunch &L, const KArgs &args, dim3 grid, hipStream_t s) {
  if (L.a_kcontig && L.b_kcontig)
    zgett_kernel<NT, BM, BN, BK, WM, WN, true, true, M3, DEEP, MINW, SEG2>
        <<<grid, NT, 0, s>>>(args);
  else if (L.a_kcontig)
    zgett_kernel<NT, BM, BN, BK, WM, WN, true, false, M3, DEEP, MINW, SEG2>
        <<<grid, NT, 0, s>>>(args);
  else if (L.b_kcontig)
    zgett_kernel<NT, BM, BN, BK, WM, WN, false, true, M3, DEEP, MINW, SEG2>
        <<<grid, NT, 0, s>>>(args);
  else
    zgett_kernel<NT, BM, BN, BK, WM, WN, false, false, M3, DEEP, MINW, SEG2>
        <<<grid, NT, 0, s>>>(args);
}

// tile shapes (complex elements): first generation 128x64; second 128x128;
// three-product kernels 128x64 with 8 waves of 32x32 (generation 3) or 4
// waves of 64x32 (generation 4); real 128x128
constexpr int kCBM = 128, kCBN = 64, kCBK = 16, kCWM = 64, kCWN = 32;
constexpr int kZBM = 128, kZBN = 128, kZBK = 16, kZWM = 64, kZWN = 32;
constexpr int kZ3BM = 128, kZ3BN = 64, kZ3BK = 16;
// short K (at most kShortK): generation 3's 32x32 wave tiles, 4 waves on a
// 64x64 tile, two workgroups per CU (66.5 KB of LDS each)
constexpr int kZSBM = 64, kZSBN = 64, kShortK = 256;
constexpr int kRBM = 128, kRBN = 128, kRBK = 16, kRWM = 64, kRWN = 64;

int g_complex_generation = 3;

bool short_k(Field f, int64_t K, bool two_segments = false) {
  return f == Field::Complex && g_complex_generation == 3 && K <= kShortK && !two_segments;
}

void tile_of(Field f, int64_t K, int &BM, int &BN, int &BK) {
  if (f == Field::Real) {
    BM = kRBM; BN = kRBN; BK = kRBK;
  } else if (g_complex_generation == 1) {
    BM = kCBM; BN = kCBN; BK = kCBK;
  } else if (short_k(f, K)) {
    BM = kZSBM; BN = kZSBN; BK = kZ3BK;
  } else if (g_complex_generation >= 3) {
    BM = kZ3BM; BN = kZ3BN; BK = kZ3BK;
  } else {
    BM = kZBM; BN = kZBN; BK = kZBK;
  }
}

inline int grid_for(int64_t n) {
  return static_cast<int>(std::min<int64_t>((n + 255) / 256, 256 * 64));
}

}  // namespace

void set_complex_gett_generation(int g) {
  g_complex_generation = (g >= 1 && g <= 5) ? g : 3;
}

std::string gett_variant(const GettLaunch &L) {
  std::string n = L.field == Field::Complex ? "gett_c64" : "gett_f64";
  n += L.a_kcontig ? "_ak" : "_am";
  n += L.b_kcontig ? "_bk" : "_bn";
  if (short_k(L.field, L.K, L.A2 != nullptr)) n += "_shortk";
  if (L.A2) n += "_seg2";
  return n;
}

void launch_gett(const GettLaunch &L, hipStream_t s, hipEvent_t start,
                 hipEvent_t stop) {
  const bool cplx = L.field == Field::Complex;
  const bool seg2 = L.A2 != nullptr;
  CC4S_ASSERT(!seg2 || (cplx && g_complex_generation == 3),
              "two-segment contractions need the complex generation-3 kernel");
  int BM, BN, BK;
  tile_of(L.field, seg2 ? kShortK + 1 : L.K, BM, BN, BK);
  KArgs a;
  a.A = static_cast<const double *>(L.A);
  a.B = static_cast<const double *>(L.B);
  a.C = static_cast<double *>(L.C);
  a.W = L.splits > 1 ? static_cast<double *>(L.W) : nullptr;
  a.rowA = L.rowA; a.colA = L.colA; a.rowB = L.rowB; a.colB = L.colB;
  a.rowC = L.rowC; a.colC = L.colC;
  a.M = L.M; a.N = L.N; a.K = L.K;
  const int64_t kstages = (L.K + BK - 1) / BK;
  a.kChunk = ((kstages + L.splits - 1) / L.splits) * BK;
  a.alpha = L.alpha; a.beta = L.beta;
  a.signA = L.conjA ? -1.0 : 1.0;
  a.signB = L.conjB ? -1.0 : 1.0;
  a.tilesM = static_cast<int>((L.M + BM - 1) / BM);
  a.tilesN = static_cast<int>((L.N + BN - 1) / BN);
  CC4S_ASSERT(static_cast<int64_t>(a.tilesM) * a.tilesN < (int64_t(1) << 31),
              "contraction too large for one launch");
  a.A2 = static_cast<const double *>(L.A2);
  a.B2 = static_cast<const double *>(L.B2);
  a.rowA2 = L.rowA2; a.colA2 = L.colA2; a.rowB2 = L.rowB2; a.colB2 = L.colB2;
  a.K1 = L.K1; a.K1p = L.K1p; a.K2 = L.K2;
  a.f2 = L.f2;
  a.signA2 = L.conjA2 ? -1.0 : 1.0;
  a.signB2 = L.conjB2 ? -1.0 : 1.0;
  dim3 grid(a.tilesM * a.tilesN, L.splits);
  if (start) CC4S_HIP(hipEventRecord(start, s));
  if (seg2)
    dispatch_zgett<512, kZ3BM, kZ3BN, kZ3BK, 32, 32, true, false, 1, true>(L, a, grid, s);
  else if (cplx && g_complex_generation == 2)
    dispatch_zgett<kZThreads, kZBM, kZBN, kZBK, kZWM, kZWN, false>(L, a, grid, s);
  else if (short_k(L.field, L.K))
    dispatch_zgett<256, kZSBM, kZSBN, kZ3BK, 32, 32, true, false, 2>(L, a, grid, s);

  else if (cplx && g_complex_generation == 3)
    dispatch_zgett<512, kZ3BM, kZ3BN, kZ3BK, 32, 32, true>(L, a, grid, s);
  else if (cplx && g_complex_generation == 4)
    dispatch_zgett<256, kZ3BM, kZ3BN, kZ3BK, 64, 32, true>(L, a, grid, s);
  else if (cplx && g_complex_generation == 5)
    dispatch_zgett<512, kZ3BM, kZ3BN, kZ3BK, 32, 32, true, true>(L, a, grid, s);
  else if (cplx)
    dispatch_gett<true, kCBM, kCBN, kCBK, kCWM, kCWN>(L, a, grid, s);
  else
    dispatch_gett<false, kRBM, kRBN, kRBK, kRWM, kRWN>(L, a, grid, s);
  CC4S_HIP(hipGetLastError());
  if (stop) CC4S_HIP(hipEventRecord(stop, s));
  if (L.splits > 1)
    launch_splitk_reduce(L.field, a.W, a.C, L.rowC, L.colC, L.M, L.N, L.splits,
                         L.alpha, L.beta, s);
}

void launch_splitk_reduce(Field f, const double *W, double *C,
                          const int64_t *rowC, const int64_t *colC, int64_t M,
                          int64_t N, int splits, double alpha, double beta,
                          hipStream_t s) {
  const int64_t total = M * N;
  if (f == Field::Complex)
    splitk_reduce_kernel<true><<<grid_for(total), 256, 0, s>>>(
        W, C, rowC, colC, M, N, splits, alpha, beta);
  else
    splitk_reduce_kernel<false><<<grid_for(total), 256, 0, s>>>(
        W, C, rowC, colC, M, N, splits, alpha, beta);
  CC4S_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// planner
// ---------------------------------------------------------------------------

struct GettPlan {
  int64_t M = 0, N = 0, K = 0;
  bool swap = false, a_kc = true, b_kc = true;
  int splits = 1;
  DeviceBuffer tables;
  const int64_t *rowA, *colA, *rowB, *colB, *rowC, *colC;
  // two-segment plans: the second contraction's tables; swap2 tells that
  // its second operand carries the M indices
  bool swap2 = false;
  int64_t K1 = 0, K1p = 0, K2 = 0;
  const int64_t *rowA2 = nullptr, *colA2 = nullptr, *rowB2 = nullptr, *colB2 = nullptr;
};

ContractionEngine::ContractionEngine(hipStream_t s) : stream_(s) { apply_env_overrides(); }

void ContractionEngine::apply_env_overrides() {
  // CC4S_MIXED_SPLIT=half|bf16 and CC4S_MIXED_MAX_K override the split
  // format and the longest FP32 sum of the mixed path
  if (const char *e = std::getenv("CC4S_MIXED_SPLIT")) {
    const std::string v = e;
    CC4S_ASSERT(v == "half" || v == "bf16", "CC4S_MIXED_SPLIT must be half or bf16");
    mixed_half = v == "half";
  }
  // CC4S_GETT_GENERATION selects the complex FP64 kernel (plans made later)
  if (const char *e = std::getenv("CC4S_GETT_GENERATION"))
    set_complex_gett_generation(std::atoi(e));
  if (const char *e = std::getenv("CC4S_MIXED_TERMS")) {
    const int v = std::atoi(e);
    CC4S_ASSERT(v == 9 || v == 12 || v == 16, "CC4S_MIXED_TERMS must be 9, 12 or 16");
    mixed_terms = v;
  }
  if (const char *e = std::getenv("CC4S_MIXED_MAX_K")) {
    const long long v = std::atoll(e);
    CC4S_ASSERT(v >= 16, "CC4S_MIXED_MAX_K must be at least 16");
    mixed_max_k = v;
  }
  // CC4S_MIXED_MIN_FLOPS: smaller complex contractions stay FP64
  if (const char *e = std::getenv("CC4S_MIXED_TILE")) set_mixed_tile(std::atoi(e));
  if (const char *e = std::getenv("CC4S_MIXED_MIN_FLOPS")) {
    const double v = std::atof(e);
    CC4S_ASSERT(v >= 0.0, "CC4S_MIXED_MIN_FLOPS must not be negative");
    mixed_min_flops = v;
  }
}

// a timing event from the pool of recorded and read ones, else a new one
hipEvent_t ContractionEngine::take_event() {
  hipEvent_t e;
  if (!event_pool_.empty()) {
    e = event_pool_.back();
    event_pool_.pop_back();
  } else {
    CC4S_HIP(hipEventCreate(&e));
  }
  return e;
}

ContractionEngine::~ContractionEngine() {
  for (auto &p : pending_) {
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
  }
  for (auto e : event_pool_) (void)hipEventDestroy(e);
}

void ContractionEngine::set_profiling(bool on) { profiling_ = on; }

void ContractionEngine::reset_timings() {
  (void)timings();
  timings_.clear();
}

const std::unordered_map<std::string, KernelTiming> &ContractionEngine::timings() {
  if (!pending_.empty()) {
    CC4S_HIP(hipStreamSynchronize(stream_));
    for (auto &p : pending_) {
      float ms = 0;
      CC4S_HIP(hipEventElapsedTime(&ms, p.start, p.stop));
      auto &t = timings_[p.name];
      t.launches += 1;
      t.ms += ms;
      t.flops += p.flops;
      event_pool_.push_back(p.start);
      event_pool_.push_back(p.stop);
    }
    pending_.clear();
  }
  return timings_;
}

namespace {

std::string plan_key(const View &A, const View &B, const View &C) {
  std::ostringstream k;
  for (const View *v : {&A, &B, &C}) {
    k << v->idx << (v->conj ? "*" : "") << ':';
    for (size_t d = 0; d < v->lens.size(); ++d)
      k << v->lens[d] << '/' << v->strides[d] << ',';
    k << '|';
  }
  k << static_cast<int>(C.field);
  return k.str();
}

// offsets of every multi-index of `group` (first listed index fastest)
std::vector<int64_t> offsets(const View &v, const std::string &group,
                             const View &shape) {
  int64_t n = 1;
  for (char c : group) n *= shape.len(c);
  std::vector<int64_t> out(static_cast<size_t>(n), 0);
  int64_t block = 1;
  for (char c : group) {
    const int64_t L = shape.len(c), s = v.stride(c);
    for (int64_t e = 0; e < n; ++e) out[e] += ((e / block) % L) * s;
    block *= L;
  }
  return out;
}

}  // namespace

GettPlan &ContractionEngine::plan_for(const View &A0, const View &B0,
                                      const View &C) {
  const std::string key = plan_key(A0, B0, C);
  auto it = plans_.find(key);
  if (it != plans_.end()) return *it->second;

  auto plan = std::make_unique<GettPlan>();
  std::string Mi, Ni, Ki;
  for (char c : C.idx) {
    const bool inA = A0.find(c) >= 0, inB = B0.find(c) >= 0;
    CC4S_ASSERT(!(inA && inB), std::string("index '") + c +
                                   "' in both operands and result is not a contraction");
    CC4S_ASSERT(inA || inB, std::string("result index '") + c + "' not in any operand");
    (inA ? Mi : Ni) += c;
  }
  for (char c : A0.idx) {
    if (C.find(c) >= 0) continue;
    CC4S_ASSERT(B0.find(c) >= 0, std::string("index '") + c + "' only in one operand");
    Ki += c;
  }
  for (char c : B0.idx)
    CC4S_ASSERT(C.find(c) >= 0 || A0.find(c) >= 0,
                std::string("index '") + c + "' only in one operand");
  for (char c : Ki)
    CC4S_ASSERT(A0.len(c) == B0.len(c), std::string("length mismatch of '") + c + "'");
  for (char c : Mi) CC4S_ASSERT(A0.len(c) == C.len(c), "length mismatch (A, C)");
  for (char c : Ni) CC4S_ASSERT(B0.len(c) == C.len(c), "length mismatch (B, C)");

  // make C contiguous along N: if its fastest index belongs to A, exchange
  // the operands and compute the transposed problem
  char cfast = 0;
  int64_t cmin = INT64_MAX;
  for (int d = 0; d < C.rank(); ++d)
    if (C.lens[d] > 1 && C.strides[d] < cmin) { cmin = C.strides[d]; cfast = C.idx[d]; }
  plan->swap = cfast && Mi.find(cfast) != std::string::npos;
  const View &A = plan->swap ? B0 : A0;
  const View &B = plan->swap ? A0 : B0;
  if (plan->swap) std::swap(Mi, Ni);

  auto by = [](const View &v) {
    return [&v](char x, char y) { return v.stride(x) < v.stride(y); };
  };
  std::stable_sort(Mi.begin(), Mi.end(), by(A));
  std::stable_sort(Ni.begin(), Ni.end(), by(C));
  bool a_unit_k = false;
  for (char c : Ki) a_unit_k |= A.stride(c) == 1 && A.len(c) > 1;
  std::stable_sort(Ki.begin(), Ki.end(), a_unit_k ? by(A) : by(B));

  auto unit = [](const View &v, const std::string &g) {
    return !g.empty() && v.stride(g[0]) == 1 && v.len(g[0]) > 1;
  };
  plan->a_kc = unit(A, Ki) || !unit(A, Mi);
  plan->b_kc = unit(B, Ki) || !unit(B, Ni);

  auto rowA = offsets(A, Mi, C), colA = offsets(A, Ki, A);
  auto rowB = offsets(B, Ki, A), colB = offsets(B, Ni, C);
  auto rowC = offsets(C, Mi, C), colC = offsets(C, Ni, C);
  plan->M = static_cast<int64_t>(rowA.size());
  plan->N = static_cast<int64_t>(colB.size());
  plan->K = static_cast<int64_t>(colA.size());
  if (Ki.empty()) plan->K = 1;  // outer product: one k with offset 0

  std::vector<int64_t> all;
  auto append = [&all](const std::vector<int64_t> &v) {
    all.insert(all.end(), v.begin(), v.end());
  };
  append(rowA); append(colA); append(rowB); append(colB); append(rowC); append(colC);
  plan->tables.allocate(all.size() * sizeof(int64_t));
  CC4S_HIP(hipMemcpy(plan->tables.get(), all.data(), all.size() * sizeof(int64_t),
                     hipMemcpyHostToDevice));
  const int64_t *base = plan->tables.as<int64_t>();
  plan->rowA = base; base += rowA.size();
  plan->colA = base; base += colA.size();
  plan->rowB = base; base += rowB.size();
  plan->colB = base; base += colB.size();
  plan->rowC = base; base += rowC.size();
  plan->colC = base;

  // split K when the tile grid cannot fill the 256 CUs twice over
  int BM, BN, BK;
  tile_of(C.field, plan->K, BM, BN, BK);
  // workgroups per CU: one for the 512-thread complex kernels, two otherwise
  const int slots = 256 * ((C.field == Field::Complex && g_complex_generation >= 2 &&
                            !short_k(C.field, plan->K)) ? 1 : 2);
  const int64_t tiles = ((plan->M + BM - 1) / BM) * ((plan->N + BN - 1) / BN);
  const int64_t kstages = (plan->K + BK - 1) / BK;
  int64_t splits = 1;
  if (tiles < 2 * slots) {
    splits = std::min<int64_t>((4 * slots + tiles - 1) / tiles, kstages / 4);
    const size_t es = scalar_bytes(C.field);
    while (splits > 1 &&
           static_cast<size_t>(splits) * plan->M * plan->N * es > max_workspace)
      --splits;
    splits = std::max<int64_t>(1, std::min<int64_t>(splits, 256));
  }
  plan->splits = static_cast<int>(splits);

  auto &ref = *plan;
  plans_.emplace(key, std::move(plan));
  return ref;
}

GettPlan &ContractionEngine::plan_for2(const View &A1, const View &B1, const View &A2,
                                       const View &B2, const View &C) {
  const std::string key = "2|" + plan_key(A1, B1, C) + "|" + plan_key(A2, B2, C);
  auto it = plans_.find(key);
  if (it != plans_.end()) return *it->second;
  // the first contraction's plan fixes the M and N orders and the swap
  const GettPlan &P1 = plan_for(A1, B1, C);
  auto plan = std::make_unique<GettPlan>();
  plan->M = P1.M;
  plan->N = P1.N;
  plan->swap = P1.swap;
  plan->a_kc = P1.a_kc;
  plan->b_kc = P1.b_kc;
  // M indices: C's indices carried by the first contraction's A-role operand,
  // in the order of its strides (as plan_for orders them)
  const View &Aop = P1.swap ? B1 : A1;
  std::string Mi, Ni;
  for (char c : C.idx) (Aop.find(c) >= 0 ? Mi : Ni) += c;
  auto by = [](const View &v) {
    return [&v](char x, char y) { return v.stride(x) < v.stride(y); };
  };
  std::stable_sort(Mi.begin(), Mi.end(), by(Aop));
  std::stable_sort(Ni.begin(), Ni.end(), by(C));
  auto c_set = [&C](const View &v) {
    std::string s;
    for (char c : v.idx)
      if (C.find(c) >= 0) s += c;
    std::sort(s.begin(), s.end());
    return s;
  };
  std::string ms = Mi, ns = Ni;
  std::sort(ms.begin(), ms.end());
  std::sort(ns.begin(), ns.end());
  CC4S_ASSERT((c_set(A2) == ms && c_set(B2) == ns) || (c_set(A2) == ns && c_set(B2) == ms),
              "the two contractions split the result indices differently");
  plan->swap2 = c_set(A2) != ms;
  const View &A = plan->swap2 ? B2 : A2;
  const View &B = plan->swap2 ? A2 : B2;
  std::string Ki;
  for (char c : A.idx) {
    if (C.find(c) >= 0) continue;
    CC4S_ASSERT(B.find(c) >= 0 && A.len(c) == B.len(c),
                std::string("index '") + c + "' of the second contraction");
    Ki += c;
  }
  CC4S_ASSERT(!Ki.empty(), "the second contraction has no summed index");
  for (char c : Mi) CC4S_ASSERT(A.len(c) == C.len(c), "length mismatch (A2, C)");
  for (char c : Ni) CC4S_ASSERT(B.len(c) == C.len(c), "length mismatch (B2, C)");
  bool a_unit_k = false;
  for (char c : Ki) a_unit_k |= A.stride(c) == 1 && A.len(c) > 1;
  std::stable_sort(Ki.begin(), Ki.end(), a_unit_k ? by(A) : by(B));
  auto rowA = offsets(A, Mi, C), colA = offsets(A, Ki, A);
  auto rowB = offsets(B, Ki, A), colB = offsets(B, Ni, C);
  CC4S_ASSERT(static_cast<int64_t>(rowA.size()) == plan->M &&
                  static_cast<int64_t>(colB.size()) == plan->N,
              "the two contractions have different result shapes");
  plan->K1 = P1.K;
  plan->K1p = (P1.K + kZ3BK - 1) / kZ3BK * kZ3BK;
  plan->K2 = static_cast<int64_t>(colA.size());
  plan->K = plan->K1p + plan->K2;
  std::vector<int64_t> all;
  all.insert(all.end(), rowA.begin(), rowA.end());
  all.insert(all.end(), colA.begin(), colA.end());
  all.insert(all.end(), rowB.begin(), rowB.end());
  all.insert(all.end(), colB.begin(), colB.end());
  plan->tables.allocate(all.size() * sizeof(int64_t));
  CC4S_HIP(hipMemcpy(plan->tables.get(), all.data(), all.size() * sizeof(int64_t),
                     hipMemcpyHostToDevice));
  const int64_t *base = plan->tables.as<int64_t>();
  plan->rowA2 = base; base += rowA.size();
  plan->colA2 = base; base += colA.size();
  plan->rowB2 = base; base += rowB.size();
  plan->colB2 = base;
  plan->rowA = P1.rowA; plan->colA = P1.colA; plan->rowB = P1.rowB; plan->colB = P1.colB;
  plan->rowC = P1.rowC; plan->colC = P1.colC;
  // split K when the tile grid cannot fill the CUs twice over (one
  // workgroup per CU)
  const int64_t tiles = ((plan->M + kZ3BM - 1) / kZ3BM) * ((plan->N + kZ3BN - 1) / kZ3BN);
  const int64_t kstages = plan->K / kZ3BK + 1;
  int64_t splits = 1;
  if (tiles < 512) {
    splits = std::min<int64_t>((1024 + tiles - 1) / tiles, kstages / 4);
    while (splits > 1 && static_cast<size_t>(splits) * plan->M * plan->N * 16 > max_workspace)
      --splits;
    splits = std::max<int64_t>(1, std::min<int64_t>(splits, 256));
  }
  plan->splits = static_cast<int>(splits);
  auto &ref = *plan;
  plans_.emplace(key, std::move(plan));
  return ref;
}

void ContractionEngine::contract2(double alpha, const View &A1, const View &B1,
                                  double alpha2, const View &A2, const View &B2,
                                  double beta, const View &C) {
  CC4S_ASSERT(A1.field == C.field && B1.field == C.field && A2.field == C.field &&
                  B2.field == C.field,
              "mixed scalar fields");
  if (C.field != Field::Complex || g_complex_generation != 3 || mixed_ || alpha == 0.0) {
    contract(alpha, A1, B1, beta, C);
    contract(alpha2, A2, B2, 1.0, C);
    return;
  }
  CC4S_ASSERT(!C.conj, "conjugated result view");
  GettPlan &P = plan_for2(A1, B1, A2, B2, C);
  if (P.M == 0 || P.N == 0) return;
  GettLaunch L;
  L.field = C.field;
  L.A = P.swap ? B1.ptr : A1.ptr;
  L.B = P.swap ? A1.ptr : B1.ptr;
  L.conjA = P.swap ? B1.conj : A1.conj;
  L.conjB = P.swap ? A1.conj : B1.conj;
  L.A2 = P.swap2 ? B2.ptr : A2.ptr;
  L.B2 = P.swap2 ? A2.ptr : B2.ptr;
  L.conjA2 = P.swap2 ? B2.conj : A2.conj;
  L.conjB2 = P.swap2 ? A2.conj : B2.conj;
  L.C = C.ptr;
  L.rowA = P.rowA; L.colA = P.colA; L.rowB = P.rowB; L.colB = P.colB;
  L.rowA2 = P.rowA2; L.colA2 = P.colA2; L.rowB2 = P.rowB2; L.colB2 = P.colB2;
  L.rowC = P.rowC; L.colC = P.colC;
  L.M = P.M; L.N = P.N; L.K = P.K;
  L.K1 = P.K1; L.K1p = P.K1p; L.K2 = P.K2;
  L.f2 = alpha2 / alpha;
  L.splits = P.splits;
  L.alpha = alpha; L.beta = beta;
  L.a_kcontig = P.a_kc; L.b_kcontig = P.b_kc;
  L.W = nullptr;
  if (P.splits > 1) {
    workspace_.reserve(static_cast<size_t>(P.splits) * P.M * P.N * 16);
    L.W = workspace_.get();
  }
  const double flops = 8.0 * double(P.M) * double(P.N) * double(P.K1 + P.K2);
  if (profiling_) {
    Pending p{take_event(), take_event(), gett_variant(L), flops};
    launch_gett(L, stream_, p.start, p.stop);
    pending_.push_back(p);
    if (pending_.size() > 4096) (void)timings();
  } else {
    launch_gett(L, stream_);
  }
  stats.flops += flops;
  stats.launches += 1;
}

void ContractionEngine::contract(double alpha, const View &A, const View &B,
                                 double beta, const View &C) {
  CC4S_ASSERT(A.field == C.field && B.field == C.field, "mixed scalar fields");
  CC4S_ASSERT(!C.conj, "conjugated result view");
  GettPlan &P = plan_for(A, B, C);
  if (P.M == 0 || P.N == 0) return;
  GettLaunch L;
  L.field = C.field;
  L.A = P.swap ? B.ptr : A.ptr;
  L.B = P.swap ? A.ptr : B.ptr;
  L.conjA = P.swap ? B.conj : A.conj;
  L.conjB = P.swap ? A.conj : B.conj;
  L.C = C.ptr;
  L.rowA = P.rowA; L.colA = P.colA; L.rowB = P.rowB; L.colB = P.colB;
  L.rowC = P.rowC; L.colC = P.colC;
  L.M = P.M; L.N = P.N; L.K = P.K;
  L.splits = P.splits;
  L.alpha = alpha; L.beta = beta;
  L.a_kcontig = P.a_kc; L.b_kcontig = P.b_kc;
  L.W = nullptr;
  if (P.splits > 1) {
    workspace_.reserve(static_cast<size_t>(P.splits) * P.M * P.N *
                       scalar_bytes(C.field));
    L.W = workspace_.get();
  }
  const double flops = (C.field == Field::Complex ? 8.0 : 2.0) * double(P.M) *
                       double(P.N) * double(P.K);
  if (mixed_ && C.field == Field::Complex && flops >= mixed_min_flops) {
    contract_mixed(P, alpha, A, B, beta, C, flops);
    stats.flops += flops;
    stats.launches += 1;
    return;
  }
  if (profiling_) {
    Pending p{take_event(), take_event(), gett_variant(L), flops};
    launch_gett(L, stream_, p.start, p.stop);
    pending_.push_back(p);
    if (pending_.size() > 4096) (void)timings();
  } else {
    launch_gett(L, stream_);
  }
  stats.flops += flops;
  stats.launches += 1;
}

void ContractionEngine::contract_mixed(GettPlan &P, double alpha, const View &A,
                                       const View &B, double beta, const View &C,
                                       double flops) {
  // 9 products per complex MAC: half parts plus a third plane of Re + Im
  const bool nine = mixed_half && mixed_terms == 9;
  const int terms = mixed_terms == 9 && !mixed_half ? 12 : mixed_terms;
  const size_t eb = nine ? 12 : 8;  // packed bytes per element and k
  const MixedTile T = mixed_tile(terms);
  const int64_t Mp = (P.M + T.BM - 1) / T.BM * T.BM;
  const int64_t Np = (P.N + T.BN - 1) / T.BN * T.BN;
  const int64_t Kfull = (P.K + T.BK - 1) / T.BK * T.BK;
  const View &Aop = P.swap ? B : A;
  const View &Bop = P.swap ? A : B;
  // an operand marked constant (a cache slot holds its pointer) is packed
  // once over the full K range
  auto slot_of = [&](const View &v) {
    for (int s = 0; s < kCacheSlots; ++s)
      if (cache_ptr_[s] && v.ptr == cache_ptr_[s] &&
          (cache_idx_[s].empty() || v.idx == cache_idx_[s]))
        return s;
    return -1;
  };
  const int slotA = slot_of(Aop);
  const int slotB = slotA < 0 ? slot_of(Bop) : -1;
  const bool cacheA = slotA >= 0, cacheB = slotB >= 0;
  const int slot = cacheA ? slotA : slotB;
  // half split: per-operand scales (A, B, the cache slots) and reduction room
  double *scA = nullptr, *scB = nullptr, *scC = nullptr, *scPart = nullptr;
  if (mixed_half) {
    scales_.reserve((2 + kCacheSlots) * sizeof(double) + split_scale_partials_bytes());
    scA = scales_.as<double>();
    scB = scA + 1;
    scC = scA + 2 + (slot > 0 ? slot : 0);
    scPart = scA + 2 + kCacheSlots;
  }
  if (cacheA || cacheB) {
    const int64_t rows = cacheA ? Mp : Np;
    if (cache_plan_[slot] != &P || cache_is_a_[slot] != cacheA || cache_eb_[slot] != eb) {
      cache_[slot].reserve(static_cast<size_t>(rows) * Kfull * eb);
      char *cs = nine ? cache_[slot].as<char>() + static_cast<size_t>(rows) * Kfull * 8
                      : nullptr;
      if (cacheA) {
        if (mixed_half)
          launch_split_scale(static_cast<const double *>(Aop.ptr), P.rowA, P.colA, P.M,
                             P.K, P.a_kc, nine, scPart, scC, stream_);
        launch_pack(static_cast<const double *>(Aop.ptr), P.rowA, P.colA, P.M, P.K, 0,
                    Kfull, Mp, Aop.conj, P.a_kc, cache_[slot].get(), cs, scC, stream_);
      } else {
        if (mixed_half)
          launch_split_scale(static_cast<const double *>(Bop.ptr), P.colB, P.rowB, P.N,
                             P.K, P.b_kc, nine, scPart, scC, stream_);
        launch_pack(static_cast<const double *>(Bop.ptr), P.colB, P.rowB, P.N, P.K, 0,
                    Kfull, Np, Bop.conj, P.b_kc, cache_[slot].get(), cs, scC, stream_);
      }
      cache_plan_[slot] = &P;
      cache_is_a_[slot] = cacheA;
      cache_eb_[slot] = eb;
    }
  }
  if (mixed_half && !cacheA)
    launch_split_scale(static_cast<const double *>(Aop.ptr), P.rowA, P.colA, P.M, P.K,
                       P.a_kc, nine, scPart, scA, stream_);
  if (mixed_half && !cacheB)
    launch_split_scale(static_cast<const double *>(Bop.ptr), P.colB, P.rowB, P.N, P.K,
                       P.b_kc, nine, scPart, scB, stream_);
  const double *useScA = mixed_half ? (cacheA ? scC : scA) : nullptr;
  const double *useScB = mixed_half ? (cacheB ? scC : scB) : nullptr;
  const int64_t packed_rows = (cacheA ? 0 : Mp) + (cacheB ? 0 : Np);
  int64_t Kc = packed_rows > 0
                   ? static_cast<int64_t>(pack_budget / (static_cast<size_t>(packed_rows) * eb))
                   : Kfull;
  Kc = std::max<int64_t>(T.BK, Kc / T.BK * T.BK);
  Kc = std::min(Kc, Kfull);
  // split K inside a chunk when the tile grid cannot fill the CUs twice
  const int64_t tiles = (Mp / T.BM) * (Np / T.BN);
  int splits = 1;
  if (tiles < 512) {
    int64_t sp = std::min<int64_t>((1024 + tiles - 1) / tiles, Kc / (4 * T.BK));
    while (sp > 1 && static_cast<size_t>(sp) * P.M * P.N * 16 > max_workspace) --sp;
    splits = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(sp, 256)));
  }
  // no FP32 sum (one split of one chunk) runs over more than mixed_max_k
  Kc = std::min<int64_t>(Kc, std::max<int64_t>(T.BK, mixed_max_k * splits / T.BK * T.BK));
  const int64_t chunks = (P.K + Kc - 1) / Kc;
  pack_.reserve(static_cast<size_t>(packed_rows) * Kc * eb);
  char *Apk = pack_.as<char>();
  char *Bpk = Apk + (cacheA ? 0 : static_cast<size_t>(Mp) * Kc * 8);
  // the Re + Im planes follow the main planes
  char *Aps = nine ? Apk + static_cast<size_t>(packed_rows) * Kc * 8 : nullptr;
  char *Bps = nine ? Aps + (cacheA ? 0 : static_cast<size_t>(Mp) * Kc * 4) : nullptr;
  const char *cacheS =
      nine && (cacheA || cacheB)
          ? cache_[slot].as<char>() + static_cast<size_t>(cacheA ? Mp : Np) * Kfull * 8
          : nullptr;
  double *W = nullptr;
  if (splits > 1) {
    workspace_.reserve(static_cast<size_t>(splits) * P.M * P.N * (nine ? 16 : 8));
    W = workspace_.as<double>();
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (profiling_) {
    e0 = take_event();
    e1 = take_event();
    CC4S_HIP(hipEventRecord(e0, stream_));
  }
  for (int64_t c = 0; c < chunks; ++c) {
    const int64_t k0 = c * Kc;
    const int64_t kc = std::min(Kc, (P.K - k0 + T.BK - 1) / T.BK * T.BK);
    const void *Ause = cacheA ? static_cast<const void *>(cache_[slot].as<char>() +
                                                          static_cast<size_t>(k0) * Mp * 8)
                              : Apk;
    const void *Buse = cacheB ? static_cast<const void *>(cache_[slot].as<char>() +
                                                          static_cast<size_t>(k0) * Np * 8)
                              : Bpk;
    const void *AuseS = !nine ? nullptr
                        : cacheA ? static_cast<const void *>(cacheS + static_cast<size_t>(k0) * Mp * 4)
                                 : Aps;
    const void *BuseS = !nine ? nullptr
                        : cacheB ? static_cast<const void *>(cacheS + static_cast<size_t>(k0) * Np * 4)
                                 : Bps;
    if (!cacheA)
      launch_pack(static_cast<const double *>(Aop.ptr), P.rowA, P.colA, P.M, P.K, k0,
                  kc, Mp, Aop.conj, P.a_kc, Apk, Aps, useScA, stream_);
    if (!cacheB)
      launch_pack(static_cast<const double *>(Bop.ptr), P.colB, P.rowB, P.N, P.K, k0,
                  kc, Np, Bop.conj, P.b_kc, Bpk, Bps, useScB, stream_);
    const double bc = c == 0 ? beta : 1.0;
    launch_bgemm(Ause, Buse, static_cast<double *>(C.ptr), W, P.rowC, P.colC, P.M,
                 P.N, Mp, Np, kc, splits, alpha, bc, useScA, useScB, terms, AuseS, BuseS,
                 stream_);
    if (splits > 1 && nine)
      launch_splitk_reduce(Field::Complex, W, static_cast<double *>(C.ptr), P.rowC,
                           P.colC, P.M, P.N, splits, alpha, bc, stream_);
    else if (splits > 1)
      launch_split_reduce(W, static_cast<double *>(C.ptr), P.rowC, P.colC, P.M, P.N,
                          splits, alpha, bc, useScA, useScB, stream_);
  }
  if (profiling_) {
    CC4S_HIP(hipEventRecord(e1, stream_));
    pending_.push_back(Pending{e0, e1, "mixed_c64", flops});
    if (pending_.size() > 4096) (void)timings();
  }
}

void ContractionEngine::sum(double alpha, const View &A, double beta,
                            const View &C) {
  CC4S_ASSERT(A.field == C.field, "mixed scalar fields");
  CC4S_ASSERT(C.rank() <= kMaxRank, "rank too large");
  CC4S_ASSERT(!C.conj, "conjugated result view");
  SumArgs a;
  a.A = static_cast<const double *>(A.ptr);
  a.C = static_cast<double *>(C.ptr);
  a.rank = C.rank();
  a.total = C.size();
  for (int d = 0; d < C.rank(); ++d) {
    const char c = C.idx[d];
    CC4S_ASSERT(A.find(c) >= 0 && A.len(c) == C.lens[d],
                std::string("sum: index '") + c + "' mismatch");
    a.lens[d] = C.lens[d];
    a.sC[d] = C.strides[d];
    a.sA[d] = A.stride(c);
  }
  CC4S_ASSERT(A.rank() == C.rank(), "sum: operand has extra indices");
  a.alpha = alpha;
  a.beta = beta;
  a.signA = A.conj ? -1.0 : 1.0;
  if (a.total == 0) return;
  if (C.field == Field::Complex)
    sum_kernel<true><<<grid_for(a.total), 256, 0, stream_>>>(a);
  else
    sum_kernel<false><<<grid_for(a.total), 256, 0, stream_>>>(a);
  CC4S_HIP(hipGetLastError());
}

void ContractionEngine::scale(double beta, const View &C) {
  SumArgs a;
  a.A = nullptr;
  a.C = static_cast<double *>(C.ptr);
  a.rank = C.rank();
  a.total = C.size();
  for (int d = 0; d < C.rank(); ++d) {
    a.lens[d] = C.lens[d];
    a.sC[d] = C.strides[d];
    a.sA[d] = 0;
  }
  a.alpha = 0.0;
  a.beta = beta;
  a.signA = 1.0;
  if (a.total == 0) return;
  if (C.field == Field::Complex)
    sum_kernel<true><<<grid_for(a.total), 256, 0, stream_>>>(a);
  else
    sum_kernel<false><<<grid_for(a.total), 256, 0, stream_>>>(a);
  CC4S_HIP(hipGetLastError());
}

}  // namespace cc4s_amd
