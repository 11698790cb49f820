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
};

// DIIS over (amplitudes, amplitude change) pairs held in HBM as ring-buffer
// slots.  The new pair is written straight into slot(next) by the
// denominator-update kernel, then append() updates the overlap matrix,
// solves the bordered system on the host in FP64 and forms the extrapolated
// amplitudes (DiisMixer.cxx:58-107).
class DiisMixer {
 public:
  DiisMixer(Field f, int64_t nset, int64_t N, hipStream_t s);
  void *next_amplitudes_slot() const { return A_[next_].get(); }
  void *next_residuum_slot() const { return R_[next_].get(); }
  // uses the pair in slot(next); writes sum_j c_j A_j to `out`
  void append(void *out);
  double residuum_norm() const { return norm_; }
  const std::vector<double> &coefficients() const { return coeffs_; }

 private:
  std::vector<double> solve(std::vector<double> matrix, size_t dim) const;
  Field f_;
  int64_t n_, N_;
  hipStream_t s_;
  std::vector<DeviceBuffer> A_, R_;
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

  CoupledClusterParams p_;
  int device_;
  hipStream_t stream_ = nullptr;
  std::unique_ptr<ContractionEngine> engine_;
  DeviceBuffer gamma_, eps_h_, eps_p_;
  std::unique_ptr<Ccsd> ccsd_;
  std::unique_ptr<DiisMixer> mixer_;
  DeviceBuffer T_, R_, scratch_, energy_partials_;
  bool have_amplitudes_ = false;
  bool inputs_set_ = false;
  bool mixed_phase_ = false;  // still iterating in mixed precision
  cplx previous_e_ = 0.0;
};

}  // namespace cc4s_amd
