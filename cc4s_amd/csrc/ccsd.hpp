// Ccsd method: CCSD residuum on one GPU, the MI355X counterpart of
// Ccsd<F,TE>::getResiduum (coupledcluster/Ccsd.cxx:76-378 real,
// 381-633 complex).  Inputs are the Coulomb vertex Gamma[G,Np,Np] resident in
// HBM and the amplitude set [T1; T2]; the output is the residuum set
// [R1; R2], all in the cc4s layout (first index fastest).
#pragma once

#include "contract.hpp"

#include <memory>

namespace cc4s_amd {

struct CcsdShape {
  Field field;
  int64_t No, Nv, NG;
  int64_t Np() const { return No + Nv; }
  int64_t n1() const { return Nv * No; }
  int64_t n2() const { return Nv * Nv * No * No; }
  int64_t nset() const { return n1() + n2(); }  // elements of one [T1;T2] set
};

enum class Method { Ccsd, Drccd };

struct CcsdOptions {
  Method method = Method::Ccsd;      // CoupledCluster.cxx:77-84
  bool ppl = true;                   // Ccsd.cxx:95,402
  int64_t integrals_slice_size = 0;  // 0: No (Ccsd.cxx:287-289)
  bool linearized = false;           // Drccd.cxx:53-55
  bool adjacent_pairs_exchange = false;  // Drccd.cxx:56-58
  // FP64 complex ladder as one fused kernel (V formed in LDS, never in HBM)
  // instead of the slab GETTs; split passes keep the slab
  bool ladder_fused = false;
  // complex ladder (FP64 and split passes): the slab contraction over pairs
  // c <= d and i <= j of symmetric and antisymmetric combinations, half the
  // multiply-adds (needs tau_cdij = tau_dcji, which the iteration keeps)
  bool ladder_paired = false;
};

// time per residuum stage, filled when profiling is on
struct CcsdTimings {
  double ladder_ms = 0, total_ms = 0;
  double ladder_flops = 0;
  int64_t ladder_calls = 0;
};

class Ccsd {
 public:
  // `gamma` is the full vertex in HBM, complex interleaved [NG, Np, Np];
  // `gamma_real` tells that every imaginary part is zero (real path only).
  Ccsd(ContractionEngine &E, const CcsdShape &shape, const CcsdOptions &opt,
       void *gamma, bool gamma_real);
  ~Ccsd();

  // Coulomb integral slices from Gamma, once per run
  // (VertexCoulombIntegrals.cxx:55-197)
  void form_integrals();

  // R = residuum(T); T == nullptr is the first iteration (R1 = 0, R2 = V)
  void get_residuum(const void *T, void *R);

  const View &Vpphh() const;

  // (T) correlation energy of the amplitudes T (PerturbativeTriplesReference.
  // cxx:55-180, cT off); eps_h, eps_p on the device.  V_ppph (No*Nv^3) is
  // formed for the call and released afterwards.
  double triples(const void *T, const double *eps_h, const double *eps_p);
  // Gamma^G_qp == conj(Gamma^G_pq) bit for bit (checked on the device)
  bool vertex_is_hermitian() const;
  CcsdTimings timings;
  bool profile = false;

 private:
  void residuum_complex(const void *T, void *R);
  void residuum_real(const void *T, void *R);
  void residuum_drccd(const void *T, void *R);
  void form_slice(Tensor &V, const char *name);
  void ladder(const View &left, const View &right, const View &tauv,
              const View &R2);
  View gam(const char *part, const std::string &names, bool conjT = false) const;

  ContractionEngine &E_;
  CcsdShape s_;
  CcsdOptions opt_;
  void *gamma_;
  bool gamma_real_;
  struct Work;
  std::unique_ptr<Work> w_;
};

}  // namespace cc4s_amd
