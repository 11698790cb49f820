// CoupledCluster driver and DiisMixer on one GPU: the MI355X counterparts
// of CoupledCluster::run (algorithms/CoupledCluster.cxx:55-317) and
// DiisMixer (mixers/DiisMixer.cxx:27-148).
#pragma once

#include "ccsd.hpp"
#include "elementwise.hpp"

#include <limits>
#include <vector>

namespace cc4s_amd {

struct CoupledClusterParams {
  CcsdShape shape;
  int64_t max_iterations = 16;           // CoupledCluster.hpp:45
  double energy_convergence = 1e-7;      // CoupledCluster.hpp:46
  double amplitudes_convergence = 1e-6;  // CoupledCluster.hpp:47
  double level_shift = 0.0;              // CoupledCluster.hpp:49
  int64_t max_residua = 4;               // DiisMixer.cxx:33
  CcsdOptions ccsd;
  // mixed precision: large contractions on split-bf16 MFMA while the
  // mixer's dR >= mixed_switch, FP64 afterwards
  bool mixed = false;
  double mixed_switch = 1e-5;
  // the mixed phase also ends after `mixed_stall_passes` passes without a
  // `mixed_stall_gain`-fold drop of dR below its best value: the split
  // residuum's rounding floor (dR ~1e-8 at No=32, Nv=256) was reached
  int mixed_stall_passes = 2;
  double mixed_stall_gain = 2.0;
  // DIIS history at the switch: its split-bf16 residua carry rounding at the
  // floor, which a late switch would extrapolate with.  0 keeps it, 1
  // restarts it, 2 restarts it when the switch comes at dR < 2e-6 or from
  // the stall guard (measured: restarting at dR ~3e-6 costs passes, at
  // dR < 1e-6 it saves them)
  int mixed_restart = 2;
  int mixed_restart_keep = 0;  // most recent pairs kept by a restart
  // what follows the mixed phase: 0 FP64 passes; 1 or 2 an FP64 anchor pass
  // that also evaluates the split residuum at the same amplitudes, giving
  // its defect d = R64(T) - Rsplit(T) (the DIIS history restarts there), then
  // split passes with Rsplit + d until they stop gaining, then FP64 passes
  // (2 restarts the DIIS history once more before them, 1 keeps it; 3 as 1,
  // but a corrected pass that meets the convergence test ends the run: the
  // defect, measured in FP64 at the anchor, stands in for the FP64 pass)
  int mixed_refine = 1;
  bool mixed_half = true;         // split parts in power-of-two scaled IEEE half
  int64_t mixed_max_k = 2048;     // longest FP32 sum of a split contraction
};

// Where the DIIS history lives: both kinds of slot in HBM; the amplitude
// slots in pinned host memory (they are read once per pass, to form the
// extrapolated amplitudes); or every slot in pinned host memory, streamed
// through HBM in chunks for the overlaps and the combination.
enum class DiisPlacement { Device, AmplitudesOnHost, AllOnHost };

// DIIS over (amplitudes, amplitude change) pairs held as ring-buffer slots
// (DiisMixer.cxx:58-107).  Slots are pair-packed sets: T1 plus the pairs
// (a,i) <= (b,j) of T2, half the size of a full set, exact because
// T2_abij = T2_baji.
class DiisMixer {
 public:
  DiisMixer(Field f, int64_t No, int64_t Nv, int64_t N, DiisPlacement where,
            hipStream_t s);
  // T' = -(R - shift T)/(Delta + shift) and dT = T' - T into slot(next)
  void write_next(const void *R, const void *T, const double *eps_h,
                  const double *eps_p, double shift);
  // overlaps, bordered solve, and `out` (full layout) = sum_j c_j A_j
  void append(void *out);
  double residuum_norm() const { return norm_; }
  // forget the stored pairs except the `keep` most recent ones, which move
  // to the first slots: the next append continues a history of keep + 1
  void restart(int keep = 0);
  const std::vector<double> &coefficients() const { return coeffs_; }
  DiisPlacement placement() const { return where_; }
  // device bytes a mixer with these parameters allocates
  static size_t device_bytes(Field f, int64_t No, int64_t Nv, int64_t N,
                             DiisPlacement where);

 private:
  std::vector<double> solve(std::vector<double> matrix, size_t dim) const;
  void append_streamed(void *out);
  Field f_;
  int64_t No_, Nv_, n_, N_;
  DiisPlacement where_;
  hipStream_t s_;
  std::vector<DeviceBuffer> A_, R_;     // HBM slots
  std::vector<HostBuffer> Ahost_, Rhost_;  // host slots
  DeviceBuffer current_;  // AllOnHost: the newest amplitude change
  DeviceBuffer staging_;  // host-slot chunks, N_ of them for AllOnHost
  int64_t chunk_ = 0;     // elements per staged chunk
  std::vector<bool> filled_;
  std::vector<double> B_;
  size_t next_ = 0, count_ = 0;
  double norm_ = 0;
  std::vector<double> coeffs_;
  DeviceBuffer partials_;
};

struct IterationReport {
  int64_t iteration = 0;
  cplx energy, direct, exchange;
  double delta_energy = 0, residuum_norm = 0, seconds = 0, flops = 0;
  bool converged = false;
  bool mixed = false;  // pass ran its large contractions in split bf16
  // 0 FP64, 1 split, 2 FP64 anchor (also evaluates the split residuum),
  // 3 split corrected by the anchor's defect
  int kind = 0;
  int diis_placement = 0;  // DiisPlacement of the mixer
  double residuum_seconds = 0, mixer_seconds = 0;
};

class CoupledCluster {
 public:
  explicit CoupledCluster(const CoupledClusterParams &p, int device);
  ~CoupledCluster();

  void set_inputs(const double *gamma, bool on_device, const double *eps);
  void set_amplitudes(const double *T1, const double *T2);
  IterationReport iterate();
  // loop + final energy; returns convergenceReached
  bool run(std::vector<IterationReport> *trace);
  void get_amplitudes(double *T1, double *T2);
  void residuum(const double *T1, const double *T2, double *R1, double *R2);
  void energy_of(const double *T1, const double *T2, cplx *E, cplx *D, cplx *X);

  // Energy adjoint at the current amplitudes.  `gamma` is a vertex
  // [NG, Np, Np] (host or device memory; null: the input vertex).
  // energy_forward: Re E of CoupledCluster.cxx:194-253 with the pphh
  // integrals formed from `gamma` (VertexCoulombIntegrals.cxx:68-171).
  double energy_forward(const double *gamma, bool gamma_on_device);
  // (T) correlation energy of the current amplitudes
  // (PerturbativeTriplesReference.cxx:55-180)
  double triples();
  // grad = dE/dRe Gamma + i dE/dIm Gamma ([NG, Np, Np] complex, host or
  // device memory, may be null) and S[G] = 1/2 sum_pq Re(conj Gamma^G_pq
  // grad^G_pq) (host, NG values, may be null); fp32: gradient products in
  // FP32 arithmetic, otherwise FP64 MFMA.
  void energy_backward(const double *gamma, bool gamma_on_device, double *grad,
                       bool grad_on_device, double *S, bool fp32);

  // Transition structure factor of FiniteSizeCorrection at the current
  // amplitudes (FiniteSizeCorrection.cxx:49-133), complex field: the vertex
  // on the momentum grid, Gamma^G = sum_F Gamma^F U[G,F] (U [NGg, NG] complex,
  // host memory; null: the vertex is on the grid already, NGg = NG), and
  // S[G] = Re sum_abij conj(Gamma^G_ai) Gamma^G_jb (2 tau_abij - tau_baij) / V[G]
  // (V [NGg] host, S [NGg] host).
  void transition_structure_factor(const double *U, int64_t NGg, const double *V,
                                   double *S);

  const CoupledClusterParams &params() const { return p_; }
  double second_order = std::numeric_limits<double>::quiet_NaN();
  cplx final_energy, final_direct, final_exchange;
  int64_t iterations_done = 0;
  Ccsd *method() { return ccsd_.get(); }
  ContractionEngine &engine() { return *engine_; }

 private:
  void energy(const void *T, cplx *E, cplx *D, cplx *X);
  void upload_set(void *dst, const double *T1, const double *T2);
  void download_set(const void *src, double *T1, double *T2);
  void create_mixer();
  const void *adjoint_vertex(const double *gamma, bool on_device);
  View vertex_block(const void *gamma, bool hp, const std::string &names) const;

  CoupledClusterParams p_;
  int device_;
  hipStream_t stream_ = nullptr;
  std::unique_ptr<ContractionEngine> engine_;
  DeviceBuffer gamma_, eps_h_, eps_p_;
  std::unique_ptr<Ccsd> ccsd_;
  std::unique_ptr<DiisMixer> mixer_;
  DeviceBuffer T_, R_, scratch_, energy_partials_;
  // energy adjoint: uploaded vertex, gradient blocks, FP32 operands, dots
  DeviceBuffer adj_gamma_, adj_hp_, adj_ph_, adj_partials_, adj_S_;
  // transition structure factor: singular vectors, grid vertex blocks, Y
  DeviceBuffer fsc_u_, fsc_ph_, fsc_hp_, fsc_y_;
  bool have_amplitudes_ = false;
  bool inputs_set_ = false;
  enum class Phase { Fp64, Mixed, Anchor, Corrected };
  Phase phase_ = Phase::Fp64;
  DeviceBuffer defect_;       // R64(T) - Rsplit(T) at the anchor pass
  double best_mixed_norm_ = 0;
  int stalled_passes_ = 0;
  // split passes end after `mixed_stall_passes` passes without a
  // `mixed_stall_gain`-fold drop of dR below their best value
  bool split_stalled(double norm);
  cplx previous_e_ = 0.0;
};

}  // namespace cc4s_amd
