// Fused particle-particle ladder on FP64 MFMA (gfx950).
//
// The reference forms V_xcyd for a pair of integral slices from the dressed
// vertices and contracts it with tau (Ccsd.cxx:579-600).  Here one kernel
// does both for every slice pair, and V exists only in LDS:
//
// * a workgroup owns 16 pairs (x, y) -- 4 x of the upper slice times 4 y of
//   the lower one -- and QN columns n = i + No j of R, and walks the (c, d)
//   plane in 16 x 16 chunks;
// * formation: V[(x,c), (y,d)] = sum_G L[G,x,c] R[G,y,d], a 64 x 64 complex
//   tile over K = G with the three-product complex MFMA, operands staged
//   through double-buffered LDS, the result stored to LDS as V[(c,d)][(x,y)];
// * contraction: acc[(x,y), n] += sum_(c,d) V[(x,y),(c,d)] tau[(c,d), n]
//   with the four-product complex MFMA (two accumulator sets, so the register
//   file holds two waves per SIMD); tau is read straight from global memory,
//   each element by one wave of the workgroup;
// * epilogue: acc is added into R2 at (a+x, b+y, i, j) and, for pairs of
//   different slices, at (b+y, a+x, j, i); every R2 element has one writer.
//
// Work per launch: (G + No^2) Nv^2 sum_pairs Nx Ny complex MACs, as on the
// slab path, except that V is formed ceil(No^2 / QN) times.
#include "ladder.hpp"

#include <algorithm>
#include <cstdlib>
#include <vector>

namespace cc4s_amd {

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int kNT = 512;             // 8 waves
constexpr int kPX = 4, kPY = 4;      // x and y of a pair block
constexpr int kPairs = kPX * kPY;    // one 16-row MFMA block
constexpr int kCh = 16;              // c (and d) per chunk
constexpr int kFM = kPX * kCh;       // formation rows (x, c)
constexpr int kFN = kPY * kCh;       // formation columns (y, d)
constexpr int kCD = kCh * kCh;       // contraction k per chunk
constexpr int kBK = 16;              // G per formation stage
constexpr int kLDF = kFM + 1;        // complex elements per LDS operand row
constexpr int kLDV = kPairs + 1;     // complex elements per LDS row of V

static_assert(kFM == kFN, "one row map serves both formation operands");
static_assert(kFM * kBK == 2 * kNT, "two elements of each operand per thread and stage");
static_assert(kPairs == 16, "the pair block is one MFMA row block");

struct LadderArgs {
  const double2 *L, *R, *tau;
  double2 *R2;
  const int4 *blocks;     // (a0, b0, nx, ny | mirror << 8)
  int64_t sLG, sLx, sLc;  // strides of L[G, x, c] in complex elements
  int64_t sRG, sRy, sRd;  // strides of R[G, y, d]
  int64_t tc, td, tn;     // tau strides of c, d and n = i + No j
  int64_t ra, rb, ri, rj; // R2 strides
  double signL, signR;    // -1 conjugates the operand
  int64_t G, v, oo;
  int o;
  int pb0, npb;           // pair blocks [pb0, pb0 + npb) of this launch
};

__device__ __forceinline__ dbl4 mfma(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// NBW: 16-column MFMA blocks per wave; QN = 8 waves * 16 * NBW columns
template <int NBW>
__global__ __launch_bounds__(kNT, 1) void ladder_fused_kernel(LadderArgs p) {
  constexpr int QN = 8 * 16 * NBW;
  __shared__ double2 Fa[2][kBK][kLDF];
  __shared__ double2 Fb[2][kBK][kLDF];
  __shared__ double2 Vt[kCD][kLDV];

  const int t = threadIdx.x;
  const int lane = t & 63, w = t >> 6;
  const int pb = p.pb0 + static_cast<int>(blockIdx.x % p.npb);
  const int64_t n0 = static_cast<int64_t>(blockIdx.x / p.npb) * QN;
  const int4 blk = p.blocks[pb];
  const int a0 = blk.x, b0 = blk.y, nx = blk.z, ny = blk.w & 0xff;
  const bool mirror = (blk.w >> 8) != 0;

  // formation operand elements of this thread: G offset fk of the stage,
  // rows fr[e] = (x or y) * 16 + (c or d) offset
  const int fk = t % kBK;
  int fr[2];
#pragma unroll
  for (int e = 0; e < 2; ++e) fr[e] = t / kBK + e * (kNT / kBK);
  int aoff[2], boff[2];  // element offsets (the host checks they fit)
  bool oka[2], okb[2];
  double2 ra[2], rb[2];
  bool gok = false;
  auto set_rows = [&](int64_t c0, int64_t d0) {
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int xy = fr[e] / kCh;
      const int64_t c = c0 + fr[e] % kCh, d = d0 + fr[e] % kCh;
      oka[e] = xy < nx && c < p.v;
      aoff[e] = oka[e] ? static_cast<int>((a0 + xy) * p.sLx + c * p.sLc) : 0;
      okb[e] = xy < ny && d < p.v;
      boff[e] = okb[e] ? static_cast<int>((b0 + xy) * p.sRy + d * p.sRd) : 0;
    }
  };
  // unconditional loads (an out-of-range element reads offset 0 and is
  // zeroed when stored), so no wait is placed right after them
  auto load_stage = [&](int s) {
    const int64_t g = static_cast<int64_t>(s) * kBK + fk;
    gok = g < p.G;
    const int64_t go = gok ? g : 0;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      ra[e] = p.L[static_cast<int64_t>(aoff[e]) + go * p.sLG];
      rb[e] = p.R[static_cast<int64_t>(boff[e]) + go * p.sRG];
    }
  };
  // zeroed component by component (a select of whole pairs would go through
  // scratch memory)
  auto store_stage = [&](int buf) {
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const bool ka = oka[e] && gok, kb = okb[e] && gok;
      Fa[buf][fk][fr[e]] = make_double2(ka ? ra[e].x : 0.0, ka ? ra[e].y * p.signL : 0.0);
      Fb[buf][fk][fr[e]] = make_double2(kb ? rb[e].x : 0.0, kb ? rb[e].y * p.signR : 0.0);
    }
  };

  // contraction columns of this lane (clamped for the loads; the epilogue
  // skips n >= No^2)
  int nIdx[NBW];
#pragma unroll
  for (int nb = 0; nb < NBW; ++nb) {
    const int64_t n = n0 + w * 16 * NBW + nb * 16 + (lane & 15);
    nIdx[nb] = static_cast<int>(n < p.oo ? n : p.oo - 1);
  }
  dbl4 accR[NBW], accI[NBW];
#pragma unroll
  for (int nb = 0; nb < NBW; ++nb) {
    accR[nb] = dbl4{0.0, 0.0, 0.0, 0.0};
    accI[nb] = dbl4{0.0, 0.0, 0.0, 0.0};
  }

  const int wr = w / 4, wc = w % 4;  // the wave's 32 x 16 part of the formation tile
  const int nst = static_cast<int>((p.G + kBK - 1) / kBK);
  const int64_t nch = (p.v + kCh - 1) / kCh;

  for (int64_t cb = 0; cb < nch; ++cb) {
    for (int64_t db = 0; db < nch; ++db) {
      const int64_t c0 = cb * kCh, d0 = db * kCh;
      // ---- formation: V[(x,c), (y,d)] over K = G ----
      set_rows(c0, d0);
      dbl4 f1[2], f2[2], f3[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        f1[i] = dbl4{0.0, 0.0, 0.0, 0.0};
        f2[i] = dbl4{0.0, 0.0, 0.0, 0.0};
        f3[i] = dbl4{0.0, 0.0, 0.0, 0.0};
      }
      // buffer 0 was last read before the barrier that preceded the
      // previous chunk's V store
      load_stage(0);
      store_stage(0);
      __syncthreads();
      for (int s = 0; s < nst; ++s) {
        const int buf = s & 1;
        if (s + 1 < nst) load_stage(s + 1);
#pragma unroll
        for (int kq = 0; kq < kBK / 4; ++kq) {
          const int kr = kq * 4 + (lane >> 4);
          double2 a[2];
#pragma unroll
          for (int i = 0; i < 2; ++i) a[i] = Fa[buf][kr][wr * 32 + i * 16 + (lane & 15)];
          const double2 b = Fb[buf][kr][wc * 16 + (lane & 15)];
#pragma unroll
          for (int i = 0; i < 2; ++i) f1[i] = mfma(a[i].x, b.x, f1[i]);
#pragma unroll
          for (int i = 0; i < 2; ++i) f2[i] = mfma(a[i].y, b.y, f2[i]);
          const double bs = b.x + b.y;
#pragma unroll
          for (int i = 0; i < 2; ++i) f3[i] = mfma(a[i].x + a[i].y, bs, f3[i]);
        }
        // the other buffer was read in stage s-1, released by its barrier
        if (s + 1 < nst) store_stage(buf ^ 1);
        __syncthreads();
      }
      // V to LDS as Vt[(c,d)][(x,y)]: every wave has passed this chunk's
      // formation barriers, so the previous chunk's contraction is done.
      // Rows (x, c) and columns (y, d) out of range hold zeros.
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int cc = 4 * r + (lane >> 4), dd = lane & 15;
          const double p1 = f1[i][r], p2 = f2[i][r];
          Vt[cc + kCh * dd][(wr * 2 + i) + kPX * wc] =
              make_double2(p1 - p2, f3[i][r] - (p1 + p2));
        }
      __syncthreads();

      // ---- contraction: acc[(x,y), n] += V[(x,y), (c,d)] tau[(c,d), n] ----
      // lane group lane >> 4 takes k = 4 kq + (lane >> 4) = c + 16 d
      double2 bA[NBW], bB[NBW];
      auto load_b = [&](double2 (&b)[NBW], int kq) {
        const int k = kq * 4 + (lane >> 4);
        const int64_t c = min(c0 + (k % kCh), p.v - 1);
        const int64_t d = min(d0 + (k / kCh), p.v - 1);
        const double2 *base = p.tau + c * p.tc + d * p.td;
#pragma unroll
        for (int nb = 0; nb < NBW; ++nb) b[nb] = base[static_cast<int64_t>(nIdx[nb]) * p.tn];
      };
      auto mac = [&](const double2 (&b)[NBW], int kq) {
        const double2 a = Vt[kq * 4 + (lane >> 4)][lane & 15];
#pragma unroll
        for (int nb = 0; nb < NBW; ++nb) accR[nb] = mfma(a.x, b[nb].x, accR[nb]);
#pragma unroll
        for (int nb = 0; nb < NBW; ++nb) accI[nb] = mfma(a.x, b[nb].y, accI[nb]);
        const double nai = -a.y;
#pragma unroll
        for (int nb = 0; nb < NBW; ++nb) accR[nb] = mfma(nai, b[nb].y, accR[nb]);
#pragma unroll
        for (int nb = 0; nb < NBW; ++nb) accI[nb] = mfma(a.y, b[nb].x, accI[nb]);
      };
      // tau for step kq+1 is in flight while step kq multiplies (two
      // register sets, indexed by constants)
      load_b(bA, 0);
      for (int kq = 0; kq < kCD / 4; kq += 2) {
        load_b(bB, kq + 1);
        mac(bA, kq);
        if (kq + 2 < kCD / 4) load_b(bA, kq + 2);
        mac(bB, kq + 1);
      }
    }
  }

  // ---- epilogue: R2 += acc, and the mirror block for different slices ----
#pragma unroll
  for (int nb = 0; nb < NBW; ++nb) {
    const int64_t n = n0 + w * 16 * NBW + nb * 16 + (lane & 15);
    if (n >= p.oo) continue;
    const int64_t i = n % p.o, j = n / p.o;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int pr = 4 * r + (lane >> 4);
      const int x = pr % kPX, y = pr / kPX;
      if (x >= nx || y >= ny) continue;
      const double vr = accR[nb][r], vi = accI[nb][r];
      double2 *q = p.R2 + (a0 + x) * p.ra + (b0 + y) * p.rb + i * p.ri + j * p.rj;
      double2 old = *q;
      *q = make_double2(old.x + vr, old.y + vi);
      if (mirror) {
        q = p.R2 + (b0 + y) * p.ra + (a0 + x) * p.rb + j * p.ri + i * p.rj;
        old = *q;
        *q = make_double2(old.x + vr, old.y + vi);
      }
    }
  }
}

}  // namespace

double ladder_fused(const View &left, const View &right, const View &tau,
                    const View &R2, int64_t slice, FusedLadderTable &table,
                    hipStream_t s) {
  CC4S_ASSERT(left.field == Field::Complex && right.field == Field::Complex &&
                  tau.field == Field::Complex && R2.field == Field::Complex,
              "fused ladder: complex field only");
  CC4S_ASSERT(left.rank() == 3 && right.rank() == 3 && tau.rank() == 4 && R2.rank() == 4,
              "fused ladder: operand ranks");
  const int64_t G = left.lens[0], v = left.lens[1], o = tau.lens[2];
  CC4S_ASSERT(left.lens[2] == v && right.lens[0] == G && right.lens[1] == v &&
                  right.lens[2] == v,
              "fused ladder: vertex shapes");
  CC4S_ASSERT(tau.lens[0] == v && tau.lens[1] == v && tau.lens[3] == o && !tau.conj &&
                  tau.strides[3] == o * tau.strides[2],
              "fused ladder: tau must be [Nv, Nv, No, No] with (i, j) dense");
  CC4S_ASSERT(R2.lens[0] == v && R2.lens[1] == v && R2.lens[2] == o && R2.lens[3] == o &&
                  !R2.conj,
              "fused ladder: R2 must be [Nv, Nv, No, No]");
  CC4S_ASSERT(slice >= 1, "fused ladder: slice size");
  const int64_t lim = int64_t(1) << 31;
  CC4S_ASSERT(v * (std::abs(left.strides[1]) + std::abs(left.strides[2])) < lim &&
                  v * (std::abs(right.strides[1]) + std::abs(right.strides[2])) < lim &&
                  left.strides[1] >= 0 && left.strides[2] >= 0 && right.strides[1] >= 0 &&
                  right.strides[2] >= 0,
              "fused ladder: vertex offsets beyond 32 bits");
  const int64_t oo = o * o;
  CC4S_ASSERT(oo < (int64_t(1) << 31), "fused ladder: No^2 too large");
  if (G == 0 || v == 0 || o == 0) return 0.0;

  // pair blocks in the reference's slice-pair order (m <= n, Ccsd.cxx:579-600),
  // built once per (Nv, slice size)
  double pairs = 0;
  const int64_t ns = (v + slice - 1) / slice;
  std::vector<int4> tbl;
  for (int64_t m = 0; m < ns; ++m)
    for (int64_t n = m; n < ns; ++n) {
      const int64_t a = n * slice, b = m * slice;
      const int64_t Nx = std::min(slice, v - a), Ny = std::min(slice, v - b);
      pairs += static_cast<double>(Nx) * Ny;
      if (table.v == v && table.slice == slice) continue;
      for (int64_t x0 = 0; x0 < Nx; x0 += kPX)
        for (int64_t y0 = 0; y0 < Ny; y0 += kPY) {
          const int nx = static_cast<int>(std::min<int64_t>(kPX, Nx - x0));
          const int ny = static_cast<int>(std::min<int64_t>(kPY, Ny - y0));
          tbl.push_back(make_int4(static_cast<int>(a + x0), static_cast<int>(b + y0), nx,
                                  ny | ((n > m ? 1 : 0) << 8)));
        }
    }
  if (table.v != v || table.slice != slice) {
    const size_t bytes = tbl.size() * sizeof(int4);
    table.blocks.reserve(bytes);
    CC4S_HIP(hipMemcpyAsync(table.blocks.get(), tbl.data(), bytes, hipMemcpyHostToDevice, s));
    CC4S_HIP(hipStreamSynchronize(s));  // tbl is pageable host memory
    table.v = v;
    table.slice = slice;
    table.count = static_cast<int>(tbl.size());
  }
  const int nblocks = table.count;

  LadderArgs a{};
  a.L = static_cast<const double2 *>(left.ptr);
  a.R = static_cast<const double2 *>(right.ptr);
  a.tau = static_cast<const double2 *>(tau.ptr);
  a.R2 = static_cast<double2 *>(R2.ptr);
  a.blocks = static_cast<const int4 *>(table.blocks.get());
  a.sLG = left.strides[0];
  a.sLx = left.strides[1];
  a.sLc = left.strides[2];
  a.sRG = right.strides[0];
  a.sRy = right.strides[1];
  a.sRd = right.strides[2];
  a.tc = tau.strides[0];
  a.td = tau.strides[1];
  a.tn = tau.strides[2];
  a.ra = R2.strides[0];
  a.rb = R2.strides[1];
  a.ri = R2.strides[2];
  a.rj = R2.strides[3];
  a.signL = left.conj ? -1.0 : 1.0;
  a.signR = right.conj ? -1.0 : 1.0;
  a.G = G;
  a.v = v;
  a.oo = oo;
  a.o = static_cast<int>(o);
  // launches of at most 2048 workgroups (8 per CU), so that no launch runs
  // for more than a few seconds at the largest sizes
  const bool wide = oo > 256;
  const int64_t nq = wide ? (oo + 1023) / 1024 : 1;
  const int per_launch = static_cast<int>(std::max<int64_t>(1, 2048 / nq));
  for (int pb0 = 0; pb0 < nblocks; pb0 += per_launch) {
    a.pb0 = pb0;
    a.npb = std::min(per_launch, nblocks - pb0);
    const dim3 grid(static_cast<unsigned>(a.npb * nq));
    if (wide)
      ladder_fused_kernel<8><<<grid, dim3(kNT), 0, s>>>(a);
    else
      ladder_fused_kernel<2><<<grid, dim3(kNT), 0, s>>>(a);
    CC4S_HIP(hipGetLastError());
  }
  return 8.0 * static_cast<double>(G + oo) * static_cast<double>(v) * v * pairs;
}

}  // namespace cc4s_amd
