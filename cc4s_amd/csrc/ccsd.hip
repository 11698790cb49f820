// CCSD residuum on the GPU.  Each statement below is one tcc statement of
// Ccsd.cxx (cited by line); three-factor products are split into two
// pairwise contractions in the cheaper order.  Conjugate transposes and
// slices of Gamma are strided views (no copies); the particle-particle
// ladder forms V_xycd slabs from the dressed vertices (or, as an option,
// forms V tiles in LDS inside one fused kernel) and never holds the full v^4
// tensor.
#include "ccsd.hpp"
#include "elementwise.hpp"
#include "ladder.hpp"
#include "triples.hpp"

#include <algorithm>
#include <chrono>
#include <cstdlib>

namespace cc4s_amd {

namespace {

inline void *offset_ptr(const void *p, int64_t elements, Field f) {
  return static_cast<char *>(const_cast<void *>(p)) + elements * scalar_bytes(f);
}

}  // namespace

struct Ccsd::Work {
  // Coulomb integral slices kept in HBM: pphh, hhhh, hhhp (both fields),
  // phhh and hhpp (complex).  For an exactly Hermitian vertex hhpp is the
  // conjugate transpose of pphh (V_klcd = conj V_cdkl) and is not stored.
  // phhp and phph are never formed: their three uses go through Gamma.
  Tensor Vpphh, Vhhpp, Vhhhh, Vhhhp, Vphhh;
  View Vhhpp_view;
  // intermediates; Xakci lives in Xakic's buffer (Xakic is consumed first)
  Tensor X, Y, Kac, Lac, Kki, Lki, Kck, Xakic, Xklij;
  View Xakci;
  Tensor DGph, DGpp, DGhh, Dpp;  // dressed vertices
  Tensor wG, Zgvo, Wgvo, Wvooo, Woo, Zgoo;
  Tensor slab, Rxy, Rpm;  // Rpm: the paired ladder's Rp and Rm of a chunk
  FusedLadderTable fused;  // pair blocks of the fused ladder
  int64_t NGe = 0;  // length of the G index of the real-path vertex views
  bool hermitian = false;
};

Ccsd::Ccsd(ContractionEngine &E, const CcsdShape &shape, const CcsdOptions &opt,
           void *gamma, bool gamma_real)
    : E_(E), s_(shape), opt_(opt), gamma_(gamma), gamma_real_(gamma_real),
      w_(new Work) {
  if (opt_.integrals_slice_size <= 0) opt_.integrals_slice_size = s_.No;
  const Field f = s_.field;
  const int64_t o = s_.No, v = s_.Nv;
  // the real path sees Re and Im of Gamma as one real index of length 2 NG
  // (Ccsd.cxx:122-136 splits them; every product there is Re*Re + Im*Im)
  w_->NGe = f == Field::Real ? (gamma_real_ ? s_.NG : 2 * s_.NG) : s_.NG;
  const int64_t G = w_->NGe;
  auto &w = *w_;
  w.X.reset(f, "abij", {v, v, o, o});
  w.Kac.reset(f, "ac", {v, v});
  w.Lac.reset(f, "ac", {v, v});
  w.Kki.reset(f, "ki", {o, o});
  w.Lki.reset(f, "ki", {o, o});
  w.Kck.reset(f, "ck", {v, o});
  w.Xakic.reset(f, "akic", {v, o, o, v});
  w.Xakci = dense_view(w.Xakic.buf.get(), f, "akci", {v, o, v, o});
  w.Xklij.reset(f, "klij", {o, o, o, o});
  w.DGph.reset(f, "Gai", {G, v, o});
  w.DGpp.reset(f, "Gab", {G, v, v});
  w.DGhh.reset(f, "Gij", {G, o, o});
  w.wG.reset(f, "G", {G});
  w.Zgvo.reset(f, "Gak", {G, v, o});
  w.Wgvo.reset(f, "Gai", {G, v, o});
  w.Wvooo.reset(f, "akij", {v, o, o, o});
  w.Woo.reset(f, "ki", {o, o});
  if (f == Field::Complex) {
    w.Y.reset(f, "abij", {v, v, o, o});
    w.Zgoo.reset(f, "Gkj", {G, o, o});
  } else if (opt_.method == Method::Drccd && opt_.adjacent_pairs_exchange) {
    w.Y.reset(f, "abij", {v, v, o, o});
  }
}

Ccsd::~Ccsd() = default;

const View &Ccsd::Vpphh() const { return w_->Vpphh.v; }

// Gamma slice `part` (two of h/p) with its dimensions named `names`;
// conjT gives the conjugate transpose cT Gamma_XY[G,x,y] = conj Gamma_YX[G,y,x]
// (Ccsd.cxx:431-436).  Real path: the real view over (Re, Im) pairs.
View Ccsd::gam(const char *part, const std::string &names, bool conjT) const {
  const int64_t No = s_.No, Np = s_.Np(), NG = s_.NG;
  auto range = [&](char c) {
    return c == 'h' ? std::make_pair<int64_t, int64_t>(0, int64_t(No))
                    : std::make_pair<int64_t, int64_t>(int64_t(No), int64_t(Np));
  };
  View full;
  if (s_.field == Field::Complex) {
    full = dense_view(gamma_, Field::Complex, "GPQ", {NG, Np, Np});
  } else {
    full.ptr = gamma_;
    full.field = Field::Real;
    full.idx = "GPQ";
    full.lens = {w_->NGe, Np, Np};
    full.strides = {gamma_real_ ? 2 : 1, 2 * NG, 2 * NG * Np};
  }
  View v;
  if (!conjT) {
    auto r1 = range(part[0]), r2 = range(part[1]);
    v = full.sliced(1, r1.first, r1.second).sliced(2, r2.first, r2.second);
  } else {
    auto r1 = range(part[1]), r2 = range(part[0]);
    v = full.sliced(1, r1.first, r1.second).sliced(2, r2.first, r2.second);
    v = v.transposed({0, 2, 1});
    if (s_.field == Field::Complex) v = v.conjugated();
  }
  return v.as(names);
}

// Coulomb integral slice `name` (four of h/p, LO RO LI RI) into V
void Ccsd::form_slice(Tensor &V, const char *name) {
  const Field f = s_.field;
  const int64_t o = s_.No, v = s_.Nv;
  auto len = [&](char c) { return c == 'h' ? o : v; };
  V.reset(f, "pqsr", {len(name[0]), len(name[1]), len(name[2]), len(name[3])});
  const std::string LI{name[0], name[2]}, RI{name[1], name[3]};
  if (f == Field::Complex) {
    // V[p,q,s,r] = sum_G conj(Gamma_{LI LO}[G,s,p]) Gamma_{RO RI}[G,q,r]
    // (VertexCoulombIntegrals.cxx:143-171)
    E_.contract(1.0, gam(LI.c_str(), "Gps", true), gam(RI.c_str(), "Gqr"), 0.0,
                V("pqsr"));
  } else {
    // V[p,q,s,r] = sum_G ReGamma_{LO LI}[G,p,s] ReGamma_{RO RI}[G,q,r] + Im Im
    // (VertexCoulombIntegrals.cxx:87-102)
    E_.contract(1.0, gam(LI.c_str(), "Gps"), gam(RI.c_str(), "Gqr"), 0.0,
                V("pqsr"));
  }
}

void Ccsd::form_integrals() {
  auto &w = *w_;
  const Field f = s_.field;
  const int64_t v = s_.Nv;
  auto make = [&](Tensor &V, const char *name) { form_slice(V, name); };
  make(w.Vpphh, "pphh");
  make(w.Vhhhh, "hhhh");
  make(w.Vhhhp, "hhhp");
  if (f == Field::Complex) {
    make(w.Vphhh, "phhh");
    w.hermitian = is_hermitian(E_.stream(), gamma_, s_.NG, s_.Np());
    if (w.hermitian) {
      // conj(V_cdkl) = sum_G Gamma[G,k,c] conj Gamma[G,d,l]
      //             = sum_G conj(Gamma[G,c,k]) Gamma[G,l,d] = V_klcd
      w.Vhhpp_view = w.Vpphh.v.transposed({2, 3, 0, 1}).conjugated();
    } else {
      make(w.Vhhpp, "hhpp");
      w.Vhhpp_view = w.Vhhpp.v;
      // the two dressed pp vertices differ only for a non-Hermitian vertex
      w.Dpp.reset(f, "Gab", {s_.NG, v, v});
    }
  }
  CC4S_HIP(hipStreamSynchronize(E_.stream()));
}

bool Ccsd::vertex_is_hermitian() const { return w_->hermitian; }

void Ccsd::get_residuum(const void *T, void *R) {
  auto t0 = std::chrono::steady_clock::now();
  if (opt_.method == Method::Drccd)
    residuum_drccd(T, R);
  else if (s_.field == Field::Complex)
    residuum_complex(T, R);
  else
    residuum_real(T, R);
  if (profile) {
    CC4S_HIP(hipStreamSynchronize(E_.stream()));
    timings.total_ms += std::chrono::duration<double, std::milli>(
                            std::chrono::steady_clock::now() - t0).count();
  }
}

// Particle-particle ladder (Ccsd.cxx:549-600 complex, 283-333 real):
// for slice pairs m <= n of size s, V_xycd = sum_G left[G,x,c] right[G,y,d]
// (x in slice n, y in slice m), R_xyij = V_xycd tau_cdij added to R2 at
// (a,b) and, for a > b, transposed at (b,a,j,i).  The x range of a slice is
// processed in chunks so that a V slab stays within a fixed HBM budget.
// With ladder_fused (FP64 complex passes) one kernel forms V in LDS and
// contracts it, and no slab exists.
void Ccsd::ladder(const View &left, const View &right, const View &tauv,
                  const View &R2) {
  auto &w = *w_;
  const Field f = s_.field;
  const int64_t v = s_.Nv, o = s_.No, s = opt_.integrals_slice_size;
  const int64_t ns = (v + s - 1) / s;
  const size_t es = scalar_bytes(f);
  const bool fused = opt_.ladder_fused && f == Field::Complex && !E_.mixed();
  // paired: tau's combinations go to Y's buffer (v^2 o^2 elements, unused
  // after the ring terms) and each chunk's Vp, Vm to Xakic's (free after the
  // crossed ring terms): cx * Ny * v^2 <= v^2 o^2
  const bool paired = !fused && opt_.ladder_paired && f == Field::Complex &&
                      w.Y.size() > 0 && std::min(s, v) <= o * o;
  const size_t slab_budget = size_t(2) << 30;
  const int64_t per_x = v * v * std::min(s, v) * static_cast<int64_t>(es);
  int64_t cx_max = std::max<int64_t>(1, std::min<int64_t>(s, slab_budget / per_x));
  if (paired) cx_max = std::max<int64_t>(1, std::min(cx_max, o * o / std::min(s, v)));
  if (!fused && (w.slab.size() < v * v * cx_max * std::min(s, v) ||
                 w.Rxy.size() < cx_max * std::min(s, v) * o * o)) {
    w.slab.reset(f, "cxdy", {v, cx_max, v, std::min(s, v)});
    w.Rxy.reset(f, "xyij", {cx_max, std::min(s, v), o, o});
  }
  if (paired && w.Rpm.size() < cx_max * std::min(s, v) * o * o)
    w.Rpm.reset(f, "xyn", {cx_max, std::min(s, v), o * o});
  const int64_t Kp = v * (v + 1) / 2, Km = v * (v - 1) / 2;
  const int64_t Np = o * (o + 1) / 2, Nm = o * (o - 1) / 2;
  void *Tp = nullptr, *Tm = nullptr;
  if (paired) {
    CC4S_ASSERT(tauv.strides == dense_view(tauv.ptr, f, "cdij", {v, v, o, o}).strides,
                "paired ladder: tau must be dense");
    Tp = w.Y.buf.get();
    Tm = offset_ptr(Tp, Kp * Np, f);
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (profile) {
    CC4S_HIP(hipEventCreate(&e0));
    CC4S_HIP(hipEventCreate(&e1));
    CC4S_HIP(hipEventRecord(e0, E_.stream()));
  }
  if (paired) ladder_tau_pairs(tauv.ptr, Tp, Tm, v, o, E_.stream());
  const double flops0 = E_.stats.flops;
  if (fused) {
    E_.stats.flops += ladder_fused(left.as("Gxc"), right.as("Gyd"), tauv.as("cdij"),
                                   R2.as("xyij"), s, w.fused, E_.stream());
    E_.stats.launches += 1;
  }
  // tau (or its two combinations) is read unchanged by every slab
  // contraction: in split mode it is packed once for all of them
  if (E_.mixed() && paired) {
    E_.set_operand_cache(Tp, 0, "kn");
    E_.set_operand_cache(Tm, 2, "kn");
  } else if (E_.mixed()) {
    E_.set_operand_cache(tauv.ptr);
  }
  for (int64_t m = 0; m < ns && !fused; ++m) {
    // the right-hand slice is the same for the whole row m: packed once (its
    // index names tell it from the left-hand slice of a diagonal pair, which
    // can start at the same address)
    if (E_.mixed())
      E_.set_operand_cache(right.sliced(1, m * s, m * s + std::min(s, v - m * s)).ptr, 1,
                           "Gyd");
    for (int64_t n = m; n < ns; ++n) {
      const int64_t a = n * s, b = m * s;
      const int64_t Nx = std::min(s, v - a), Ny = std::min(s, v - b);
      const View Rv = right.sliced(1, b, b + Ny).as("Gyd");
      for (int64_t x0 = 0; x0 < Nx; x0 += cx_max) {
        const int64_t cx = std::min(cx_max, Nx - x0);
        const View Lv = left.sliced(1, a + x0, a + x0 + cx).as("Gxc");
        const View slab = dense_view(w.slab.buf.get(), f, "cxdy", {v, cx, v, Ny});
        E_.contract(1.0, Lv, Rv, 0.0, slab);                       // :588-589
        const View Rxy = dense_view(w.Rxy.buf.get(), f, "xyij", {cx, Ny, o, o});
        if (paired) {                                              // :590
          void *Vp = w.Xakic.buf.get();
          void *Vm = offset_ptr(Vp, Kp * cx * Ny, f);
          void *Rp = w.Rpm.buf.get();
          void *Rm = offset_ptr(Rp, cx * Ny * Np, f);
          ladder_vertex_pairs(slab.ptr, Vp, Vm, v, cx, Ny, E_.stream());
          const double fl0 = E_.stats.flops;
          E_.contract(1.0, dense_view(Vp, f, "kxy", {Kp, cx, Ny}),
                      dense_view(Tp, f, "kn", {Kp, Np}), 0.0,
                      dense_view(Rp, f, "xyn", {cx, Ny, Np}));
          const View RmV = dense_view(Rm, f, "xyn", {cx, Ny, Nm});
          if (Nm > 0 && Km > 0)
            E_.contract(1.0, dense_view(Vm, f, "kxy", {Km, cx, Ny}),
                        dense_view(Tm, f, "kn", {Km, Nm}), 0.0, RmV);
          else if (Nm > 0)
            E_.scale(0.0, RmV);
          // credited as the slab contraction it replaces (8 flops per complex
          // MAC over Nv^2 k and No^2 columns), the convention of the others
          E_.stats.flops = fl0 + 8.0 * static_cast<double>(cx * Ny) * v * v * o * o;
          ladder_unpack_pairs(Rp, Rm, Rxy.ptr, cx * Ny, o, E_.stream());
        } else {
          E_.contract(1.0, slab, tauv.as("cdij"), 0.0, Rxy);
        }
        View blk = R2.sliced(0, a + x0, a + x0 + cx).sliced(1, b, b + Ny);
        E_.sum(1.0, Rxy, 1.0, blk.as("xyij"));                     // :591
        if (a > b) {                                               // :593-596
          View mir = R2.sliced(0, b, b + Ny).sliced(1, a + x0, a + x0 + cx);
          E_.sum(1.0, Rxy, 1.0, mir.as("yxji"));
        }
      }
    }
  }
  E_.clear_operand_cache();
  if (profile) {
    CC4S_HIP(hipEventRecord(e1, E_.stream()));
    CC4S_HIP(hipEventSynchronize(e1));
    float ms = 0;
    CC4S_HIP(hipEventElapsedTime(&ms, e0, e1));
    timings.ladder_ms += ms;
    timings.ladder_flops += E_.stats.flops - flops0;
    timings.ladder_calls += 1;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
}

// Drccd::getResiduum (Drccd.cxx:31-109), both fields.  The phhp and hhpp
// integrals are contracted through Gamma:
//   2 V_akic T_cbkj = 2 sum_G L[G,a,i] Y[G,b,j],
//       Y[G,b,j] = sum_kc Gamma_hp[G,k,c] T_cbkj;
//   2 W_klcd T_acik T_dblj with W = 2 V: 4 sum_G A[G,a,i] Y[G,b,j],
//       A[G,a,i] = sum_kc M[G,k,c] T_acik;
// complex: L[G,a,i] = conj Gamma_hp[G,i,a], M[G,k,c] = conj Gamma_ph[G,c,k]
// (V[p,q,s,r] = sum_G conj(Gamma[G,s,p]) Gamma[G,q,r]); real: L = Gamma_ph,
// M = Gamma_hp over the Re/Im index (VertexCoulombIntegrals.cxx:68-102).
// With adjacentPairsExchange the -V_lkcd part of W does not factorise and is
// contracted from the explicit integrals.
void Ccsd::residuum_drccd(const void *Tp, void *Rp) {
  auto &w = *w_;
  auto &E = E_;
  const Field f = s_.field;
  const bool cplx = f == Field::Complex;
  const int64_t o = s_.No, v = s_.Nv;
  const hipStream_t st = E.stream();
  const View R1 = dense_view(Rp, f, "ai", {v, o});
  const View R2 = dense_view(offset_ptr(Rp, s_.n1(), f), f, "abij", {v, v, o, o});
  E.scale(0.0, R1);                                               // :61,76
  if (!Tp) {                                                      // :62
    E.sum(1.0, w.Vpphh("abij"), 0.0, R2);
    return;
  }
  const View T2 = dense_view(offset_ptr(Tp, s_.n1(), f), f, "abij", {v, v, o, o});
  const View Yg = w.Wgvo("Gbj");
  E.contract(1.0, gam("hp", "Gkc"), T2.as("cbkj"), 0.0, Yg);
  E.contract(2.0, gam("ph", "Gai", cplx), Yg, 0.0, R2);          // :77
  symmetrize_pairs(st, f, R2.ptr, o, v);                          // :78
  E.sum(1.0, w.Vpphh("abij"), 1.0, R2);                           // :79
  if (opt_.linearized) return;                                    // :80-83
  const View Ag = w.Zgvo("Gai");
  E.contract(1.0, gam("hp", "Gkc", cplx), T2.as("acik"), 0.0, Ag);
  E.contract(4.0, Ag, Yg, 1.0, R2);                               // :97-98
  if (opt_.adjacent_pairs_exchange) {                             // :88-91
    View Vlkcd;
    if (cplx) {
      Vlkcd = w.Vhhpp_view.as("lkcd");
    } else {
      // real V_klcd = sum_G' Gamma_hp[G',k,c] Gamma_hp[G',l,d], in X's buffer
      E.contract(1.0, gam("hp", "Gkc"), gam("hp", "Gld"), 0.0,
                 dense_view(w.X.buf.get(), f, "klcd", {o, o, v, v}));
      Vlkcd = dense_view(w.X.buf.get(), f, "lkcd", {o, o, v, v});
    }
    const View Q = dense_view(w.Y.buf.get(), f, "aldi", {v, o, v, o});
    E.contract(1.0, Vlkcd, T2.as("acik"), 0.0, Q);
    E.contract(-2.0, Q, T2.as("dblj"), 1.0, R2);
  }
}

void Ccsd::residuum_complex(const void *Tp, void *Rp) {
  auto &w = *w_;
  auto &E = E_;
  const Field f = Field::Complex;
  const int64_t o = s_.No, v = s_.Nv;
  const hipStream_t st = E.stream();
  const View R1 = dense_view(Rp, f, "ai", {v, o});
  const View R2 = dense_view(offset_ptr(Rp, s_.n1(), f), f, "abij", {v, v, o, o});
  const View Vpphh = w.Vpphh("abij");
  if (!Tp) {                                                      // :404-409
    E.scale(0.0, R1);
    E.sum(1.0, Vpphh, 0.0, R2);
    return;
  }
  const View T1 = dense_view(const_cast<void *>(Tp), f, "ai", {v, o});
  const View T2 = dense_view(offset_ptr(Tp, s_.n1(), f), f, "abij", {v, v, o, o});
  auto C = [&](double al, const View &A, const View &B, double be, const View &Cv) {
    E.contract(al, A, B, be, Cv);
  };
  auto S = [&](double al, const View &A, double be, const View &Cv) {
    E.sum(al, A, be, Cv);
  };
  auto &X = w.X, &Y = w.Y;
  auto &Vhhhp = w.Vhhhp, &Vphhh = w.Vphhh;
  auto Vhhpp = [&w](const std::string &names) { return w.Vhhpp_view.as(names); };
  auto Xakci = [&w](const std::string &names) { return w.Xakci.as(names); };

  tau(st, f, X.buf.get(), T2.ptr, T1.ptr, 1.0, o, v);            // :460-461
  tau(st, f, Y.buf.get(), T2.ptr, T1.ptr, 2.0, o, v);            // :462-463
  C(-2.0, Vhhpp("klcd"), X("adkl"), 0.0, w.Kac("ac"));          // :465
  C(1.0, Vhhpp("kldc"), X("adkl"), 1.0, w.Kac("ac"));           // :466
  S(1.0, w.Kac("ac"), 0.0, w.Lac("ac"));                        // :468
  C(1.0, gam("hp", "Gkd"), T1.as("dk"), 0.0, w.wG("G"));        // :469
  C(2.0, gam("pp", "Gac", true), w.wG("G"), 1.0, w.Lac("ac"));
  C(1.0, gam("pp", "Gad", true), T1.as("dk"), 0.0, w.Zgvo("Gak"));  // :470
  C(-1.0, w.Zgvo("Gak"), gam("hp", "Gkc"), 1.0, w.Lac("ac"));
  C(2.0, Vhhpp("klcd"), X("cdil"), 0.0, w.Kki("ki"));           // :472
  C(-1.0, Vhhpp("kldc"), X("cdil"), 1.0, w.Kki("ki"));          // :473
  S(1.0, w.Kki("ki"), 0.0, w.Lki("ki"));                        // :476
  C(2.0, Vhhhp("klic"), T1.as("cl"), 1.0, w.Lki("ki"));         // :477
  C(-1.0, Vhhhp("lkic"), T1.as("cl"), 1.0, w.Lki("ki"));        // :478
  C(1.0, w.Lac("ac"), T2.as("cbij"), 0.0, R2);                  // :481
  C(-1.0, w.Lki("ki"), T2.as("abkj"), 1.0, R2);                 // :484
  S(1.0, gam("ph", "Gai", true), 0.0, w.DGph("Gai"));           // :487
  C(-1.0, gam("hh", "Gki", true), T1.as("ak"), 1.0, w.DGph("Gai"));  // :488
  C(1.0, gam("pp", "Gbc"), T1.as("cj"), 0.0, w.Wgvo("Gbj"));    // :489
  C(1.0, w.DGph("Gai"), w.Wgvo("Gbj"), 1.0, R2);
  C(-1.0, Vphhh("akij"), T1.as("bk"), 1.0, R2);                 // :490
  // :491 with V_akic = sum_G cTGamma_ph[G,a,i] Gamma_hp[G,k,c]
  // (VertexCoulombIntegrals.cxx:166-168, slice phhp) contracted through G
  C(1.0, gam("hp", "Gkc"), T1.as("cj"), 0.0, w.Zgoo("Gkj"));
  C(1.0, gam("ph", "Gai", true), w.Zgoo("Gkj"), 0.0, w.Wvooo("akij"));
  C(-1.0, w.Wvooo("akij"), T1.as("bk"), 1.0, R2);
  C(1.0, gam("pp", "Gad", true), T1.as("di"), 1.0, w.DGph("Gai"));   // :494
  C(-0.5, gam("hp", "Gld", true), Y("dail"), 1.0, w.DGph("Gai"));   // :495
  C(1.0, w.DGph("Gai"), gam("hp", "Gkc"), 0.0, w.Xakic("akic"));    // :496
  S(1.0, Vhhpp("lkdc"), 0.0, Y("dclk"));                        // :497
  S(-0.5, Vhhpp("lkcd"), 1.0, Y("dclk"));                       // :498
  C(1.0, Y("dclk"), T2.as("adil"), 1.0, w.Xakic("akic"));       // :499
  S(2.0, T2.as("cbkj"), 0.0, Y("cbkj"));                        // :500
  S(-1.0, T2.as("bckj"), 1.0, Y("cbkj"));                       // :501
  C(1.0, w.Xakic("akic"), Y("cbkj"), 1.0, R2);                  // :502
  S(1.0, gam("pp", "Gab", true), 0.0, w.DGpp("Gab"));           // :509
  C(-1.0, gam("hp", "Glc", true), T1.as("al"), 1.0, w.DGpp("Gac"));  // :510
  S(1.0, gam("hh", "Gij"), 0.0, w.DGhh("Gij"));                 // :511
  C(1.0, gam("hp", "Gkd"), T1.as("di"), 1.0, w.DGhh("Gki"));    // :512
  C(1.0, w.DGpp("Gac"), w.DGhh("Gki"), 0.0, Xakci("akci"));     // :514
  C(-0.5, Vhhpp("lkcd"), T2.as("dail"), 1.0, Xakci("akci"));    // :515
  C(-1.0, Xakci("akci"), T2.as("cbkj"), 1.0, R2);               // :516
  C(-1.0, Xakci("bkci"), T2.as("ackj"), 1.0, R2);               // :517
  symmetrize_pairs(st, f, R2.ptr, o, v);                        // :520
  S(1.0, Vpphh, 1.0, R2);                                       // :529
  S(1.0, w.Vhhhh("klij"), 0.0, w.Xklij("klij"));                // :534
  C(1.0, Vhhhp("klic"), T1.as("cj"), 1.0, w.Xklij("klij"));     // :535
  C(1.0, Vhhhp("lkjc"), T1.as("ci"), 1.0, w.Xklij("klij"));     // :536
  C(1.0, w.Xklij("klij"), X("abkl"), 1.0, R2);                  // :538
  C(1.0, Vhhpp("klcd"), X("cdij"), 0.0, w.Xklij("klij"));       // :541
  C(1.0, w.Xklij("klij"), opt_.ppl ? T2.as("abkl") : X("abkl"), 1.0, R2);  // :542-546
  if (opt_.ppl) {
    if (w.hermitian) {
      // with Gamma_qp = conj Gamma_pq exactly, Gamma_pp[G,a,b] - Gamma_hp[G,k,b]
      // T_ak (:560-561) equals cTDGamma_pp[G,a,b] of :509-510 term by term
      ladder(w.DGpp.v, w.DGpp.v, X.v, R2);                      // :564-600
    } else {
      S(1.0, gam("pp", "Gab"), 0.0, w.Dpp("Gab"));              // :560
      C(-1.0, gam("hp", "Gkb"), T1.as("ak"), 1.0, w.Dpp("Gab"));  // :561
      ladder(w.DGpp.v, w.Dpp.v, X.v, R2);                       // :564-600
    }
  }
  C(1.0, w.Kac("ac"), T1.as("ci"), 0.0, R1);                    // :609
  C(-1.0, w.Kki("ki"), T1.as("ak"), 1.0, R1);                   // :610
  C(2.0, Vhhpp("klcd"), T1.as("dl"), 0.0, w.Kck("ck"));         // :613
  C(-1.0, Vhhpp("kldc"), T1.as("dl"), 1.0, w.Kck("ck"));        // :614
  C(2.0, w.Kck("ck"), T2.as("caki"), 1.0, R1);                  // :617
  C(-1.0, w.Kck("ck"), T2.as("caik"), 1.0, R1);                 // :618
  C(1.0, w.Kck("ck"), T1.as("ci"), 0.0, w.Woo("ki"));           // :619
  C(1.0, w.Woo("ki"), T1.as("ak"), 1.0, R1);
  // :620, V_akic T_ck = sum_G cTGamma_ph[G,a,i] (sum_kc Gamma_hp[G,k,c] T_ck)
  C(2.0, gam("ph", "Gai", true), w.wG("G"), 1.0, R1);
  // :621, V_akci = sum_G cTGamma_pp[G,a,c] Gamma_hh[G,k,i] (slice phph)
  C(1.0, gam("hh", "Gki"), T1.as("ck"), 0.0, w.Wgvo("Gci"));
  C(-1.0, gam("pp", "Gac", true), w.Wgvo("Gci"), 1.0, R1);
  C(1.0, gam("hp", "Gkd"), X("cdik"), 0.0, w.Wgvo("Gci"));      // :622-623
  C(2.0, gam("pp", "Gac", true), w.Wgvo("Gci"), 1.0, R1);
  C(1.0, gam("hp", "Gkc"), X("cdik"), 0.0, w.Wgvo("Gdi"));      // :624-625
  C(-1.0, gam("pp", "Gad", true), w.Wgvo("Gdi"), 1.0, R1);
  C(-2.0, Vhhhp("klic"), X("ackl"), 1.0, R1);                   // :627
  C(1.0, Vhhhp("lkic"), X("ackl"), 1.0, R1);                    // :628
}

void Ccsd::residuum_real(const void *Tp, void *Rp) {
  auto &w = *w_;
  auto &E = E_;
  const Field f = Field::Real;
  const int64_t o = s_.No, v = s_.Nv;
  const hipStream_t st = E.stream();
  const View R1 = dense_view(Rp, f, "ai", {v, o});
  const View R2 = dense_view(offset_ptr(Rp, s_.n1(), f), f, "abij", {v, v, o, o});
  const View Vpphh = w.Vpphh("abij");
  if (!Tp) {                                                      // :97-102
    E.scale(0.0, R1);
    E.sum(1.0, Vpphh, 0.0, R2);
    return;
  }
  const View T1 = dense_view(const_cast<void *>(Tp), f, "ai", {v, o});
  const View T2 = dense_view(offset_ptr(Tp, s_.n1(), f), f, "abij", {v, v, o, o});
  auto C = [&](double al, const View &A, const View &B, double be, const View &Cv) {
    E.contract(al, A, B, be, Cv);
  };
  auto S = [&](double al, const View &A, double be, const View &Cv) {
    E.sum(al, A, be, Cv);
  };
  auto &X = w.X;
  auto &Vp = w.Vpphh, &Vhhhp = w.Vhhhp;
  // gam(...) here is the real (Re, Im) vertex: Re*Re + Im*Im in one sum

  tau(st, f, X.buf.get(), T2.ptr, T1.ptr, 1.0, o, v);            // :169-170
  C(2.0, Vp("cdkl"), X("cdil"), 0.0, w.Kki("ki"));               // :172
  C(-1.0, Vp("dckl"), X("cdil"), 1.0, w.Kki("ki"));              // :173
  S(1.0, w.Kki("ki"), 0.0, w.Lki("ki"));                         // :175
  C(-2.0, Vp("cdkl"), X("adkl"), 0.0, w.Kac("ac"));              // :177
  C(1.0, Vp("dckl"), X("adkl"), 1.0, w.Kac("ac"));               // :178
  S(1.0, w.Kac("ac"), 0.0, w.Lac("ac"));                         // :180
  C(1.0, gam("ph", "Gdk"), T1.as("dk"), 0.0, w.wG("G"));         // :182-185
  C(2.0, gam("pp", "Gca"), w.wG("G"), 1.0, w.Lac("ac"));
  C(1.0, gam("pp", "Gda"), T1.as("dk"), 0.0, w.Zgvo("Gak"));     // :186-189
  C(-1.0, w.Zgvo("Gak"), gam("ph", "Gck"), 1.0, w.Lac("ac"));
  C(2.0, Vhhhp("klic"), T1.as("cl"), 1.0, w.Lki("ki"));          // :192
  C(-1.0, Vhhhp("lkic"), T1.as("cl"), 1.0, w.Lki("ki"));         // :193
  C(1.0, w.Lac("ac"), T2.as("cbij"), 0.0, R2);                   // :195
  C(-1.0, w.Lki("ki"), T2.as("abkj"), 1.0, R2);                  // :197
  S(1.0, gam("ph", "Gai"), 0.0, w.DGph("Gai"));                  // :201,203
  C(-1.0, gam("hh", "Gki"), T1.as("ak"), 1.0, w.DGph("Gai"));    // :202,204
  C(1.0, gam("pp", "Gbc"), T1.as("cj"), 0.0, w.Wgvo("Gbj"));     // :205-208
  C(1.0, w.DGph("Gai"), w.Wgvo("Gbj"), 1.0, R2);
  C(-1.0, Vhhhp("jika"), T1.as("bk"), 1.0, R2);                  // :209
  C(1.0, Vp("acik"), T1.as("cj"), 0.0, w.Wvooo("akij"));         // :210
  C(-1.0, w.Wvooo("akij"), T1.as("bk"), 1.0, R2);
  tau(st, f, X.buf.get(), T2.ptr, T1.ptr, 2.0, o, v);            // :215
  C(-0.5, Vp("dclk"), X("dail"), 0.0, w.Xakic("akic"));          // :216
  C(1.0, Vp("dclk"), T2.as("adil"), 1.0, w.Xakic("akic"));       // :217
  C(-0.5, Vp("cdlk"), T2.as("adil"), 1.0, w.Xakic("akic"));      // :218
  C(1.0, gam("pp", "Gad"), T1.as("di"), 1.0, w.DGph("Gai"));     // :220-221
  C(1.0, w.DGph("Gai"), gam("ph", "Gck"), 1.0, w.Xakic("akic")); // :223-224
  S(2.0, T2.as("cbkj"), 0.0, X("cbkj"));                         // :227
  S(-1.0, T2.as("bckj"), 1.0, X("cbkj"));                        // :228
  C(1.0, w.Xakic("akic"), X("cbkj"), 1.0, R2);                   // :229
  C(-0.5, Vp("cdlk"), T2.as("dail"), 0.0, w.Xakci.as("akci"));      // :239
  S(1.0, gam("pp", "Gab"), 0.0, w.DGpp("Gab"));                  // :240-241
  C(-1.0, gam("ph", "Gcl"), T1.as("al"), 1.0, w.DGpp("Gac"));    // :242-243
  S(1.0, gam("hh", "Gij"), 0.0, w.DGhh("Gij"));                  // :244-245
  C(1.0, gam("ph", "Gdk"), T1.as("di"), 1.0, w.DGhh("Gki"));     // :246-247
  C(1.0, w.DGpp("Gac"), w.DGhh("Gki"), 1.0, w.Xakci.as("akci"));    // :248-249
  C(-1.0, w.Xakci.as("akci"), T2.as("cbkj"), 1.0, R2);              // :252
  C(-1.0, w.Xakci.as("bkci"), T2.as("ackj"), 1.0, R2);              // :253
  symmetrize_pairs(st, f, R2.ptr, o, v);                         // :256
  S(1.0, Vpphh, 1.0, R2);                                        // :263
  tau(st, f, X.buf.get(), T2.ptr, T1.ptr, 1.0, o, v);            // :267-268
  S(1.0, w.Vhhhh("klij"), 0.0, w.Xklij("klij"));                 // :269
  C(1.0, Vhhhp("klic"), T1.as("cj"), 1.0, w.Xklij("klij"));      // :270
  C(1.0, Vhhhp("lkjc"), T1.as("ci"), 1.0, w.Xklij("klij"));      // :271
  C(1.0, w.Xklij("klij"), X("abkl"), 1.0, R2);                   // :273
  C(1.0, Vp("cdkl"), X("cdij"), 0.0, w.Xklij("klij"));           // :275
  C(1.0, w.Xklij("klij"), opt_.ppl ? T2.as("abkl") : X("abkl"), 1.0, R2);  // :276-280
  if (opt_.ppl) ladder(w.DGpp.v, w.DGpp.v, X.v, R2);             // :283-333
  C(1.0, w.Kac("ac"), T1.as("ci"), 0.0, R1);                     // :341
  C(-1.0, w.Kki("ki"), T1.as("ak"), 1.0, R1);                    // :342
  C(2.0, Vp("cdkl"), T1.as("dl"), 0.0, w.Kck("ck"));             // :345
  C(-1.0, Vp("dckl"), T1.as("dl"), 1.0, w.Kck("ck"));            // :346
  C(2.0, w.Kck("ck"), T2.as("caki"), 1.0, R1);                   // :349
  C(-1.0, w.Kck("ck"), T2.as("caik"), 1.0, R1);                  // :350
  C(1.0, w.Kck("ck"), T1.as("ci"), 0.0, w.Woo("ki"));            // :351
  C(1.0, w.Woo("ki"), T1.as("ak"), 1.0, R1);
  C(2.0, Vp("acik"), T1.as("ck"), 1.0, R1);                      // :352
  C(1.0, gam("pp", "Gca"), T1.as("ck"), 0.0, w.Zgvo("Gak"));     // :358-361
  C(-1.0, w.Zgvo("Gak"), gam("hh", "Gik"), 1.0, R1);
  C(1.0, gam("ph", "Gdk"), X("cdik"), 0.0, w.Wgvo("Gci"));       // :363-366
  C(2.0, gam("pp", "Gca"), w.Wgvo("Gci"), 1.0, R1);
  C(1.0, gam("ph", "Gck"), X("cdik"), 0.0, w.Wgvo("Gdi"));       // :367-370
  C(-1.0, gam("pp", "Gda"), w.Wgvo("Gdi"), 1.0, R1);
  C(-2.0, Vhhhp("klic"), X("ackl"), 1.0, R1);                    // :372
  C(1.0, Vhhhp("lkic"), X("ackl"), 1.0, R1);                     // :373
}

}  // namespace cc4s_amd

namespace cc4s_amd {

namespace {

// `v` with index `c` fixed at `value` (the dimension dropped)
View fixed(const View &v, char c, int64_t value) {
  const int d = v.find(c);
  CC4S_ASSERT(d >= 0 && value >= 0 && value < v.lens[d], "fixed: bad index");
  View r = v;
  r.ptr = static_cast<char *>(v.ptr) + value * v.strides[d] * static_cast<int64_t>(scalar_bytes(v.field));
  r.idx.erase(d, 1);
  r.lens.erase(r.lens.begin() + d);
  r.strides.erase(r.strides.begin() + d);
  return r;
}

}  // namespace

// (T): for every hole pair (i, j) and chunk of k, the connected amplitudes
// W_abcijk = sum over the six pair permutations of
//   T'_{x1x2x3,y1y2y3} = V_{x2 x3 d y3} T_{x1 d y1 y2} - conj(V_{y2 y3 l x3}) T_{x1 x2 y1 l}
// (PerturbativeTriplesReference.cxx:106-116) are accumulated by twelve
// contractions into W[a,b,c,k], each reading its operands in place through
// index-permuted views (no transposes); the energy kernel then adds the
// disconnected part and the denominator (:118-163).  Only i <= j <= k is
// formed (the energy kernel adds the other orderings through the symmetry of
// W), so the particle contractions cost No^3 Nv^4 MACs in total instead of
// 6 No^3 Nv^4.
double Ccsd::triples(const void *Tp, const double *eps_h, const double *eps_p) {
  auto &w = *w_;
  auto &E = E_;
  const Field f = s_.field;
  const bool cplx = f == Field::Complex;
  const int64_t o = s_.No, v = s_.Nv;
  const size_t es = scalar_bytes(f);
  Tensor Vppph;
  form_slice(Vppph, "ppph");                                     // V[b,c,d,k]
  const View T1 = dense_view(const_cast<void *>(Tp), f, "ai", {v, o});
  const View T2 = dense_view(offset_ptr(Tp, s_.n1(), f), f, "abij", {v, v, o, o});
  auto t2 = [&](const char *names) { return T2.as(names); };
  auto vp = [&](const char *names) { return Vppph.v.as(names); };
  const View Vh0 = cplx ? w.Vhhhp.v.conjugated() : w.Vhhhp.v;   // conj V_hhhp
  auto vh = [&](const char *names) { return Vh0.as(names); };

  // bytes of W per hole-triple chunk (CC4S_TRIPLES_W_BYTES overrides)
  size_t budget = size_t(4) << 30;
  if (const char *e = std::getenv("CC4S_TRIPLES_W_BYTES")) budget = std::strtoull(e, nullptr, 10);
  const int64_t per_k = v * v * v * static_cast<int64_t>(es);
  const int64_t kc = std::max<int64_t>(1, std::min<int64_t>(o, static_cast<int64_t>(budget) / per_k));
  Tensor Wt(f, "abck", {v, v, v, kc});
  DeviceBuffer acc(kTriplesBlocks * sizeof(double));
  CC4S_HIP(hipMemsetAsync(acc.get(), 0, kTriplesBlocks * sizeof(double), E.stream()));
  for (int64_t i = 0; i < o; ++i) {
    for (int64_t j = i; j < o; ++j) {
      for (int64_t k0 = j; k0 < o; k0 += kc) {
        const int64_t kn = std::min(kc, o - k0);
        const View W = Wt.v.sliced(3, 0, kn);
        auto K = [&](const View &x) { return x.sliced(x.find('k'), k0, k0 + kn); };
        // the six particle terms T'_{abc,ijk}, T'_{bac,jik}, T'_{acb,ikj},
        // T'_{cba,kji}, T'_{cab,kij}, T'_{bca,jki} (K = d) and the six hole
        // terms (K = l), each hole term paired with the particle term that
        // splits W's indices the same way ({a}|{bck}, {b}|{ack}, {ak}|{bc},
        // {ck}|{ab} twice, {bk}|{ac}): one launch per pair, W read and
        // written six times instead of twelve
        E.contract2(1.0, fixed(fixed(t2("adij"), 'i', i), 'j', j), K(vp("bcdk")),
                    -1.0, K(t2("cbkl")), fixed(fixed(vh("jila"), 'j', j), 'i', i), 0.0, W);
        E.contract2(1.0, fixed(fixed(t2("bdji"), 'j', j), 'i', i), K(vp("acdk")),
                    -1.0, K(t2("cakl")), fixed(fixed(vh("ijlb"), 'i', i), 'j', j), 1.0, W);
        E.contract2(1.0, K(fixed(t2("adik"), 'i', i)), fixed(vp("cbdj"), 'j', j),
                    -1.0, fixed(t2("bcjl"), 'j', j), K(fixed(vh("kila"), 'i', i)), 1.0, W);
        E.contract2(1.0, K(fixed(t2("cdkj"), 'j', j)), fixed(vp("badi"), 'i', i),
                    -1.0, fixed(t2("abil"), 'i', i), K(fixed(vh("jklc"), 'j', j)), 1.0, W);
        E.contract2(1.0, K(fixed(t2("cdki"), 'i', i)), fixed(vp("abdj"), 'j', j),
                    -1.0, fixed(t2("bajl"), 'j', j), K(fixed(vh("iklc"), 'i', i)), 1.0, W);
        E.contract2(1.0, K(fixed(t2("bdjk"), 'j', j)), fixed(vp("cadi"), 'i', i),
                    -1.0, fixed(t2("acil"), 'i', i), K(fixed(vh("kjlb"), 'j', j)), 1.0, W);
        triples_energy(E.stream(), f, W.ptr, w.Vpphh.v.ptr, T1.ptr, eps_h, eps_p, o, v, i, j,
                       k0, kn, acc.as<double>());
      }
    }
  }
  std::vector<double> h(kTriplesBlocks);
  CC4S_HIP(hipMemcpyAsync(h.data(), acc.get(), kTriplesBlocks * sizeof(double),
                          hipMemcpyDeviceToHost, E.stream()));
  CC4S_HIP(hipStreamSynchronize(E.stream()));
  double e = 0.0;
  for (double x : h) e += x;
  return e;
}

}  // namespace cc4s_amd
