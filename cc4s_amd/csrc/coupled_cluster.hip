#include "coupled_cluster.hpp"

#include "adjoint.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>

namespace cc4s_amd {

// ---------------------------------------------------------------------------
// DiisMixer
// ---------------------------------------------------------------------------

namespace {
constexpr size_t kStagingBytes = size_t(1) << 30;   // amplitude-slot transfer chunk
constexpr size_t kSlotChunkBytes = size_t(1) << 29; // per-slot chunk, all on host
}

size_t DiisMixer::device_bytes(Field f, int64_t No, int64_t Nv, int64_t N,
                               DiisPlacement where) {
  const size_t set = static_cast<size_t>(packed_set_size(No, Nv)) * scalar_bytes(f);
  switch (where) {
    case DiisPlacement::Device: return 2 * N * set;
    case DiisPlacement::AmplitudesOnHost: return N * set + std::min(set, kStagingBytes);
    case DiisPlacement::AllOnHost: return set + N * std::min(set, kSlotChunkBytes);
  }
  return 0;
}

DiisMixer::DiisMixer(Field f, int64_t No, int64_t Nv, int64_t N, DiisPlacement where,
                     hipStream_t s)
    : f_(f), No_(No), Nv_(Nv), n_(packed_set_size(No, Nv)), N_(N), where_(where), s_(s) {
  CC4S_ASSERT(N >= 1 && N <= 16, "maxResidua must be in [1, 16]");
  const size_t es = scalar_bytes(f);
  const size_t set = static_cast<size_t>(n_) * es;
  // host-slot transfer chunk (CC4S_DIIS_CHUNK_BYTES lowers it)
  size_t chunk_cap = where_ == DiisPlacement::AllOnHost ? kSlotChunkBytes : kStagingBytes;
  if (const char *c = std::getenv("CC4S_DIIS_CHUNK_BYTES")) {
    const long long v = std::atoll(c);
    CC4S_ASSERT(v >= static_cast<long long>(es), "CC4S_DIIS_CHUNK_BYTES too small");
    chunk_cap = std::min(chunk_cap, static_cast<size_t>(v));
  }
  // CC4S_DIIS_PINNED=0 keeps host slots in pageable memory
  const char *pe = std::getenv("CC4S_DIIS_PINNED");
  const bool pin = !(pe && std::string(pe) == "0");
  switch (where_) {
    case DiisPlacement::Device:
      A_.resize(N);
      R_.resize(N);
      for (int64_t i = 0; i < N; ++i) {
        A_[i].allocate(set);
        R_[i].allocate(set);
      }
      break;
    case DiisPlacement::AmplitudesOnHost:
      R_.resize(N);
      for (int64_t i = 0; i < N; ++i) R_[i].allocate(set);
      for (int64_t i = 0; i < N; ++i) Ahost_.emplace_back(set, pin);
      chunk_ = static_cast<int64_t>(std::min(set, chunk_cap) / es);
      staging_.allocate(chunk_ * es);
      break;
    case DiisPlacement::AllOnHost:
      // the change slots first: they are read twice per pass (overlaps and
      // norm), so they get pinned memory while the system grants it
      for (int64_t i = 0; i < N; ++i) Rhost_.emplace_back(set, pin);
      for (int64_t i = 0; i < N; ++i) Ahost_.emplace_back(set, pin);
      current_.allocate(set);
      chunk_ = static_cast<int64_t>(std::min(set, chunk_cap) / es);
      staging_.allocate(static_cast<size_t>(N) * chunk_ * es);
      break;
  }
  filled_.assign(N, false);
  const size_t M = N + 1;
  B_.assign(M * M, 0.0);                         // DiisMixer.cxx:42-44
  for (size_t i = 1; i < M; ++i) B_[i] = -1.0;
  for (size_t i = M; i < M * M; i += M) B_[i] = -1.0;
  partials_.allocate(std::max(dots_partials_size(static_cast<int>(N)),
                              dots_partials_size(1)));
}

// B c = (-1, 0, ..., 0) for the symmetric bordered DIIS matrix (the
// reference calls LAPACK dsysv/zsysv, DiisMixer.cxx:119-148); Gaussian
// elimination with partial pivoting in FP64 gives the same solution for a
// nonsingular system.  An exactly singular system fails like info != 0.
std::vector<double> DiisMixer::solve(std::vector<double> a, size_t n) const {
  std::vector<double> x(n, 0.0);
  x[0] = -1.0;
  for (size_t k = 0; k < n; ++k) {
    size_t piv = k;
    for (size_t r = k + 1; r < n; ++r)
      if (std::fabs(a[r * n + k]) > std::fabs(a[piv * n + k])) piv = r;
    if (a[piv * n + k] == 0.0) throw Error("Diagonalization failed");
    if (piv != k) {
      for (size_t c = 0; c < n; ++c) std::swap(a[k * n + c], a[piv * n + c]);
      std::swap(x[k], x[piv]);
    }
    for (size_t r = k + 1; r < n; ++r) {
      const double l = a[r * n + k] / a[k * n + k];
      if (l == 0.0) continue;
      for (size_t c = k; c < n; ++c) a[r * n + c] -= l * a[k * n + c];
      x[r] -= l * x[k];
    }
  }
  for (size_t k = n; k-- > 0;) {
    double s = x[k];
    for (size_t c = k + 1; c < n; ++c) s -= a[k * n + c] * x[c];
    x[k] = s / a[k * n + k];
  }
  return x;
}

void DiisMixer::write_next(const void *R, const void *T, const double *eps_h,
                           const double *eps_p, double shift) {
  const size_t set = static_cast<size_t>(n_) * scalar_bytes(f_);
  if (where_ == DiisPlacement::Device) {
    amplitude_update_packed(s_, f_, R, T, eps_h, eps_p, shift, A_[next_].get(),
                            R_[next_].get(), No_, Nv_);
    return;
  }
  // T' goes through a device set to the host: the update kernel runs for T'
  // alone, the set is copied out, then it runs again for dT (on the first
  // pass dT = T' is already in place)
  void *dev = where_ == DiisPlacement::AllOnHost ? current_.get() : R_[next_].get();
  amplitude_update_packed(s_, f_, R, T, eps_h, eps_p, shift, dev, nullptr, No_, Nv_);
  CC4S_HIP(hipMemcpyAsync(Ahost_[next_].get(), dev, set, hipMemcpyDeviceToHost, s_));
  if (T)
    amplitude_update_packed(s_, f_, R, T, eps_h, eps_p, shift, nullptr, dev, No_, Nv_);
  if (where_ == DiisPlacement::AllOnHost)
    CC4S_HIP(hipMemcpyAsync(Rhost_[next_].get(), dev, set, hipMemcpyDeviceToHost, s_));
  CC4S_HIP(hipStreamSynchronize(s_));
}

void DiisMixer::restart(int keep) {
  const size_t N = static_cast<size_t>(N_);
  const size_t k = std::min<size_t>(static_cast<size_t>(std::max(keep, 0)), count_);
  // slots of the k most recent pairs, oldest first
  std::vector<size_t> src(k);
  for (size_t a = 0; a < k; ++a) src[a] = (next_ + N - k + a) % N;
  // kept pairs to slots 0..k-1, every other buffer after them
  std::vector<bool> used(N, false);
  std::vector<size_t> order;
  for (size_t a = 0; a < k; ++a) {
    order.push_back(src[a]);
    used[src[a]] = true;
  }
  for (size_t i = 0; i < N; ++i)
    if (!used[i]) order.push_back(i);
  auto permute = [&](auto &slots) {
    if (slots.empty()) return;
    std::remove_reference_t<decltype(slots)> old(N);
    for (size_t i = 0; i < N; ++i) old[i] = std::move(slots[i]);
    for (size_t i = 0; i < N; ++i) slots[i] = std::move(old[order[i]]);
  };
  permute(A_);
  permute(R_);
  permute(Ahost_);
  permute(Rhost_);
  const std::vector<double> B = B_;
  for (size_t a = 0; a < k; ++a)
    for (size_t b = 0; b < k; ++b)
      B_[(a + 1) * (N + 1) + b + 1] = B[(src[a] + 1) * (N + 1) + src[b] + 1];
  filled_.assign(N, false);
  for (size_t a = 0; a < k; ++a) filled_[a] = true;
  count_ = k;
  next_ = k % N;
}

void DiisMixer::append(void *out) {
  const size_t N = static_cast<size_t>(N_);
  const size_t es = scalar_bytes(f_);
  filled_[next_] = true;                          // DiisMixer.cxx:65-66
  // overlaps 2 Re <R_i|R_new> with every stored residuum (:69-77)
  std::vector<size_t> which;
  for (size_t i = 0; i < N; ++i)
    if (filled_[i]) which.push_back(i);
  const int nw = static_cast<int>(which.size());
  std::vector<cplx> dots(nw, cplx(0.0, 0.0));
  if (where_ != DiisPlacement::AllOnHost) {
    std::vector<const void *> xs;
    for (size_t i : which) xs.push_back(R_[i].get());
    multi_dot_packed(s_, f_, xs.data(), nw, R_[next_].get(), No_, Nv_,
                     partials_.get(), dots);
  } else {
    // every stored change streams through its staging chunk; the newest is
    // resident.  Chunk sums are added in a fixed order.
    std::vector<const void *> xs(nw);
    std::vector<cplx> part;
    for (int64_t t0 = 0; t0 < n_; t0 += chunk_) {
      const int64_t len = std::min(chunk_, n_ - t0);
      const char *y = static_cast<const char *>(current_.get()) + t0 * es;
      for (int k = 0; k < nw; ++k) {
        if (which[k] == next_) {
          xs[k] = y;
          continue;
        }
        char *dst = static_cast<char *>(staging_.get()) + k * chunk_ * es;
        CC4S_HIP(hipMemcpyAsync(dst, static_cast<const char *>(Rhost_[which[k]].get()) + t0 * es,
                                len * es, hipMemcpyHostToDevice, s_));
        xs[k] = dst;
      }
      multi_dot_packed_range(s_, f_, xs.data(), nw, y, t0, len, No_, Nv_,
                             partials_.get(), part);
      for (int k = 0; k < nw; ++k) dots[k] += part[k];
    }
  }
  for (int k = 0; k < nw; ++k) {
    const size_t i = which[k];
    const double overlap = 2.0 * dots[k].real();
    B_[(i + 1) * (N + 1) + next_ + 1] = overlap;
    B_[(next_ + 1) * (N + 1) + i + 1] = overlap;
  }
  if (count_ < N) ++count_;                       // :80
  const size_t dim = count_ + 1;
  std::vector<double> matrix(dim * dim);
  for (size_t m = 0; m < dim; ++m)
    for (size_t n = 0; n < dim; ++n) matrix[n + dim * m] = B_[n + (N + 1) * m];
  const std::vector<double> column = solve(matrix, dim);   // :89
  // next = sum_j c_j A_(next-j), residuum = sum_j c_j R_(next-j) (:92-104)
  std::vector<size_t> slots;
  coeffs_.clear();
  for (size_t j = 0; j < count_; ++j) {
    const size_t i = (next_ + N - j) % N;
    slots.push_back(i);
    coeffs_.push_back(column[i + 1]);
  }
  double rr = 0.0;
  if (where_ == DiisPlacement::Device) {
    std::vector<const void *> As, Rs;
    for (size_t i : slots) {
      As.push_back(A_[i].get());
      Rs.push_back(R_[i].get());
    }
    rr = combine_packed(s_, f_, As.data(), Rs.data(), coeffs_, No_, Nv_, out,
                        partials_.get());
  } else {
    // stream every host amplitude slot through the staging buffer,
    // accumulating in the same order (most recent first)
    for (size_t j = 0; j < slots.size(); ++j) {
      const char *src = static_cast<const char *>(Ahost_[slots[j]].get());
      for (int64_t t0 = 0; t0 < n_; t0 += chunk_) {
        const int64_t len = std::min(chunk_, n_ - t0);
        CC4S_HIP(hipMemcpyAsync(staging_.get(), src + t0 * es, len * es,
                                hipMemcpyHostToDevice, s_));
        unpack_accumulate(s_, f_, staging_.get(), t0, len, coeffs_[j], j == 0, out,
                          No_, Nv_);
      }
    }
    if (where_ == DiisPlacement::AmplitudesOnHost) {
      std::vector<const void *> Rs;
      for (size_t i : slots) Rs.push_back(R_[i].get());
      rr = norm_packed(s_, f_, Rs.data(), coeffs_, No_, Nv_, partials_.get());
    } else {
      std::vector<const void *> Rs(slots.size());
      for (int64_t t0 = 0; t0 < n_; t0 += chunk_) {
        const int64_t len = std::min(chunk_, n_ - t0);
        for (size_t j = 0; j < slots.size(); ++j) {
          if (slots[j] == next_) {
            Rs[j] = static_cast<const char *>(current_.get()) + t0 * es;
            continue;
          }
          char *dst = static_cast<char *>(staging_.get()) + j * chunk_ * es;
          CC4S_HIP(hipMemcpyAsync(dst, static_cast<const char *>(Rhost_[slots[j]].get()) + t0 * es,
                                  len * es, hipMemcpyHostToDevice, s_));
          Rs[j] = dst;
        }
        rr += norm_packed_range(s_, f_, Rs.data(), coeffs_, t0, len, No_, Nv_,
                                partials_.get());
      }
    }
  }
  norm_ = std::sqrt(rr);                          // :106
  next_ = (next_ + 1) % N;                        // :105
}

// ---------------------------------------------------------------------------
// CoupledCluster
// ---------------------------------------------------------------------------

CoupledCluster::CoupledCluster(const CoupledClusterParams &p, int device)
    : p_(p), device_(device) {
  const auto &s = p_.shape;
  CC4S_ASSERT(s.No > 0 && s.Nv > 0 && s.NG > 0, "empty dimension");
  // a residuum with a single occupied orbital made an illegal device memory
  // access whose kernel is not identified; such contexts are refused
  CC4S_ASSERT(s.No >= 2, "No = 1 (a single occupied orbital) is not supported");
  CC4S_HIP(hipSetDevice(device_));
  CC4S_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  engine_.reset(new ContractionEngine(stream_));
  const size_t es = scalar_bytes(s.field);
  T_.allocate(s.nset() * es);
  R_.allocate(s.nset() * es);
  energy_partials_.allocate(energy_partials_size());
  CC4S_ASSERT(p_.max_residua >= 1 && p_.max_residua <= 16,
              "maxResidua must be in [1, 16]");
  CC4S_ASSERT(!p_.mixed || s.field == Field::Complex,
              "mixed precision needs the complex field");
  phase_ = p_.mixed ? Phase::Mixed : Phase::Fp64;
  engine_->mixed_half = p_.mixed_half;
  engine_->mixed_max_k = p_.mixed_max_k;
  engine_->apply_env_overrides();
}

// The DIIS history is the largest allocation; it is placed once the vertex,
// the integrals and the intermediates are resident: in HBM if it fits, else
// with the amplitude slots in pinned host memory, else with every slot there
// (override with CC4S_DIIS_PLACEMENT=device|host|host_all).
void CoupledCluster::create_mixer() {
  const auto &s = p_.shape;
  size_t free_b = 0, total_b = 0;
  CC4S_HIP(hipMemGetInfo(&free_b, &total_b));
  // room kept for buffers allocated on first use (ladder slab and block,
  // operand packs, split-K partials; in mixed mode the packed tau of the
  // ladder, Nv^2 x No^2 rounded to the kernel tile, 8 bytes each)
  size_t reserve = size_t(6) << 30;
  if (p_.mixed) {
    const int64_t rows = (s.No * s.No + 127) / 128 * 128;
    reserve += static_cast<size_t>(s.Nv) * s.Nv * rows * 8;
    // the defect of the split residuum (one pair-packed set)
    if (p_.mixed_refine)
      reserve += static_cast<size_t>(packed_set_size(s.No, s.Nv)) * scalar_bytes(s.field);
  }
  const size_t avail = free_b > reserve ? free_b - reserve : 0;
  const char *env = std::getenv("CC4S_DIIS_PLACEMENT");
  const std::string placement = env ? env : "auto";
  auto need = [&](DiisPlacement w) {
    return DiisMixer::device_bytes(s.field, s.No, s.Nv, p_.max_residua, w);
  };
  DiisPlacement where;
  if (placement == "device") {
    where = DiisPlacement::Device;
  } else if (placement == "host") {
    where = DiisPlacement::AmplitudesOnHost;
  } else if (placement == "host_all") {
    where = DiisPlacement::AllOnHost;
  } else {
    CC4S_ASSERT(placement == "auto",
                "CC4S_DIIS_PLACEMENT must be auto, device, host or host_all");
    if (need(DiisPlacement::Device) <= avail)
      where = DiisPlacement::Device;
    else if (need(DiisPlacement::AmplitudesOnHost) <= avail)
      where = DiisPlacement::AmplitudesOnHost;
    else
      where = DiisPlacement::AllOnHost;
    CC4S_ASSERT(need(where) <= avail, "out of device memory for the DIIS history");
  }
  mixer_.reset(new DiisMixer(s.field, s.No, s.Nv, p_.max_residua, where, stream_));
}

CoupledCluster::~CoupledCluster() {
  ccsd_.reset();
  mixer_.reset();
  engine_.reset();
  if (stream_) (void)hipStreamDestroy(stream_);
}

void CoupledCluster::set_inputs(const double *gamma, bool on_device,
                                const double *eps) {
  CC4S_HIP(hipSetDevice(device_));
  const auto &s = p_.shape;
  const size_t gbytes = static_cast<size_t>(s.NG) * s.Np() * s.Np() * 16;
  gamma_.allocate(gbytes);
  bool gamma_real = false;
  if (on_device) {
    CC4S_HIP(hipMemcpyAsync(gamma_.get(), gamma, gbytes, hipMemcpyDeviceToDevice, stream_));
  } else {
    CC4S_HIP(hipMemcpyAsync(gamma_.get(), gamma, gbytes, hipMemcpyHostToDevice, stream_));
    if (s.field == Field::Real) {
      gamma_real = true;
      const size_t n = gbytes / 16;
      for (size_t e = 0; e < n && gamma_real; ++e) gamma_real = gamma[2 * e + 1] == 0.0;
    }
  }
  eps_h_.allocate(s.No * sizeof(double));
  eps_p_.allocate(s.Nv * sizeof(double));
  CC4S_HIP(hipMemcpyAsync(eps_h_.get(), eps, s.No * sizeof(double),
                          hipMemcpyHostToDevice, stream_));
  CC4S_HIP(hipMemcpyAsync(eps_p_.get(), eps + s.No, s.Nv * sizeof(double),
                          hipMemcpyHostToDevice, stream_));
  ccsd_.reset(new Ccsd(*engine_, s, p_.ccsd, gamma_.get(), gamma_real));
  ccsd_->form_integrals();
  CC4S_HIP(hipStreamSynchronize(stream_));
  if (!mixer_) create_mixer();
  inputs_set_ = true;
}

void CoupledCluster::upload_set(void *dst, const double *T1, const double *T2) {
  const auto &s = p_.shape;
  const size_t es = scalar_bytes(s.field);
  CC4S_HIP(hipMemcpyAsync(dst, T1, s.n1() * es, hipMemcpyHostToDevice, stream_));
  CC4S_HIP(hipMemcpyAsync(static_cast<char *>(dst) + s.n1() * es, T2, s.n2() * es,
                          hipMemcpyHostToDevice, stream_));
  CC4S_HIP(hipStreamSynchronize(stream_));
}

void CoupledCluster::download_set(const void *src, double *T1, double *T2) {
  const auto &s = p_.shape;
  const size_t es = scalar_bytes(s.field);
  if (T1)
    CC4S_HIP(hipMemcpyAsync(T1, src, s.n1() * es, hipMemcpyDeviceToHost, stream_));
  if (T2)
    CC4S_HIP(hipMemcpyAsync(T2, static_cast<const char *>(src) + s.n1() * es,
                            s.n2() * es, hipMemcpyDeviceToHost, stream_));
  CC4S_HIP(hipStreamSynchronize(stream_));
}

void CoupledCluster::set_amplitudes(const double *T1, const double *T2) {
  CC4S_HIP(hipSetDevice(device_));
  upload_set(T_.get(), T1, T2);
  have_amplitudes_ = true;
}

void CoupledCluster::energy(const void *T, cplx *E, cplx *D, cplx *X) {
  const auto &s = p_.shape;
  cc4s_amd::energy(stream_, s.field, T, ccsd_->Vpphh().ptr, s.No, s.Nv,
                   energy_partials_.get(), D, X);
  *E = *D + *X;
}

bool CoupledCluster::split_stalled(double norm) {
  if (best_mixed_norm_ == 0 || norm * p_.mixed_stall_gain < best_mixed_norm_) {
    best_mixed_norm_ = norm;
    stalled_passes_ = 0;
  } else {
    ++stalled_passes_;
  }
  return stalled_passes_ >= p_.mixed_stall_passes;
}

IterationReport CoupledCluster::iterate() {
  CC4S_ASSERT(inputs_set_, "inputs not set");
  CC4S_HIP(hipSetDevice(device_));
  const auto &s = p_.shape;
  const auto t0 = std::chrono::steady_clock::now();
  const double flops0 = engine_->stats.flops;
  IterationReport rep;
  rep.iteration = iterations_done + 1;
  const void *T = have_amplitudes_ ? T_.get() : nullptr;
  if (phase_ == Phase::Anchor && T) {
    // the split residuum and the FP64 residuum at the same amplitudes:
    // d = R64(T) - Rsplit(T), kept pair-packed (both residua satisfy
    // R_abij = R_baji, and the amplitude update reads one element of each
    // pair); the pass itself continues with R64 and a DIIS history that
    // holds no split residua
    defect_.reserve(packed_set_size(s.No, s.Nv) * scalar_bytes(s.field));
    engine_->set_mixed(true);
    ccsd_->get_residuum(T, R_.get());
    engine_->set_mixed(false);
    pack_axpby(stream_, s.field, R_.get(), -1.0, defect_.get(), 0.0, s.No, s.Nv);
    ccsd_->get_residuum(T, R_.get());                   // :136
    pack_axpby(stream_, s.field, R_.get(), 1.0, defect_.get(), 1.0, s.No, s.Nv);
    mixer_->restart(0);
    rep.mixed = false;
    rep.kind = 2;
  } else {
    rep.mixed = phase_ == Phase::Mixed || phase_ == Phase::Corrected;
    rep.kind = phase_ == Phase::Mixed ? 1 : phase_ == Phase::Corrected ? 3 : 0;
    engine_->set_mixed(rep.mixed);
    ccsd_->get_residuum(T, R_.get());                   // :136
    engine_->set_mixed(false);
    if (phase_ == Phase::Corrected)
      unpack_accumulate(stream_, s.field, defect_.get(), 0, packed_set_size(s.No, s.Nv),
                        1.0, false, R_.get(), s.No, s.Nv);
  }
  CC4S_HIP(hipStreamSynchronize(stream_));
  const auto t1 = std::chrono::steady_clock::now();
  // :137-145: T' = -R/(Delta+s) and dT = T' - T, written into the mixer slot
  mixer_->write_next(R_.get(), T, static_cast<const double *>(eps_h_.get()),
                     static_cast<const double *>(eps_p_.get()), p_.level_shift);
  const bool is_second_order = !have_amplitudes_;
  mixer_->append(T_.get());                             // :145-147
  have_amplitudes_ = true;
  CC4S_HIP(hipStreamSynchronize(stream_));
  const auto t2 = std::chrono::steady_clock::now();
  rep.residuum_seconds = std::chrono::duration<double>(t1 - t0).count();
  rep.mixer_seconds = std::chrono::duration<double>(t2 - t1).count();
  rep.diis_placement = static_cast<int>(mixer_->placement());
  rep.residuum_norm = mixer_->residuum_norm();          // :148
  energy(T_.get(), &rep.energy, &rep.direct, &rep.exchange);   // :149
  CC4S_HIP(hipStreamSynchronize(stream_));
  rep.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  rep.flops = engine_->stats.flops - flops0;
  rep.delta_energy = (rep.energy - previous_e_).real();
  if (is_second_order) second_order = rep.energy.real();      // :164-166
  rep.converged = std::abs(rep.energy - previous_e_) < p_.energy_convergence &&
                  rep.residuum_norm < p_.amplitudes_convergence;  // :167-173
  // the convergence test only counts FP64 passes; the mixed phase ends once
  // the extrapolated amplitude change falls below the switch threshold, or
  // stops falling (the rounding floor of the split residuum)
  if (rep.mixed) rep.converged = false;
  switch (phase_) {
    case Phase::Fp64:
      break;
    case Phase::Mixed:
      if (!is_second_order) {
        const bool stalled = split_stalled(rep.residuum_norm);
        if (rep.residuum_norm < p_.mixed_switch || stalled) {
          if (p_.mixed_refine) {
            phase_ = Phase::Anchor;
          } else {
            phase_ = Phase::Fp64;
            if (p_.mixed_restart == 1 ||
                (p_.mixed_restart == 2 && (stalled || rep.residuum_norm < 2e-6)))
              mixer_->restart(p_.mixed_restart_keep);
          }
        }
      }
      break;
    case Phase::Anchor:
      phase_ = Phase::Corrected;
      best_mixed_norm_ = rep.residuum_norm;
      stalled_passes_ = 0;
      break;
    case Phase::Corrected: {
      // FP64 passes take over once a corrected pass meets the convergence
      // test (an FP64 pass then has to confirm it) or the passes stall at
      // the noise of the split residuum
      const bool meets = std::abs(rep.energy - previous_e_) < p_.energy_convergence &&
                         rep.residuum_norm < p_.amplitudes_convergence;
      if (meets && p_.mixed_refine == 3) {
        rep.converged = true;
        break;
      }
      if (split_stalled(rep.residuum_norm) || meets) {
        phase_ = Phase::Fp64;
        if (p_.mixed_refine == 2) mixer_->restart(0);
      }
      break;
    }
  }
  previous_e_ = rep.energy;
  iterations_done += 1;
  return rep;
}

bool CoupledCluster::run(std::vector<IterationReport> *trace) {
  int64_t i = 0;
  for (; i < p_.max_iterations; ++i) {                  // :129
    IterationReport r = iterate();
    if (trace) trace->push_back(r);
    if (r.converged) break;
  }
  CC4S_ASSERT(have_amplitudes_, "no amplitudes: maxIterations 0 needs initialAmplitudes");
  energy(T_.get(), &final_energy, &final_direct, &final_exchange);   // :184
  return i < p_.max_iterations;                         // :185
}

void CoupledCluster::get_amplitudes(double *T1, double *T2) {
  CC4S_ASSERT(have_amplitudes_, "no amplitudes yet");
  CC4S_HIP(hipSetDevice(device_));
  download_set(T_.get(), T1, T2);
}

void CoupledCluster::residuum(const double *T1, const double *T2, double *R1,
                              double *R2) {
  CC4S_ASSERT(inputs_set_, "inputs not set");
  CC4S_HIP(hipSetDevice(device_));
  const auto &s = p_.shape;
  const void *T = nullptr;
  if (T1 && T2) {
    scratch_.reserve(s.nset() * scalar_bytes(s.field));
    upload_set(scratch_.get(), T1, T2);
    T = scratch_.get();
  }
  // a context created for mixed precision evaluates its residuum that way
  engine_->set_mixed(p_.mixed);
  ccsd_->get_residuum(T, R_.get());
  engine_->set_mixed(false);
  download_set(R_.get(), R1, R2);
}

void CoupledCluster::energy_of(const double *T1, const double *T2, cplx *E,
                               cplx *D, cplx *X) {
  CC4S_ASSERT(inputs_set_, "inputs not set");
  CC4S_HIP(hipSetDevice(device_));
  const auto &s = p_.shape;
  scratch_.reserve(s.nset() * scalar_bytes(s.field));
  upload_set(scratch_.get(), T1, T2);
  energy(scratch_.get(), E, D, X);
}


double CoupledCluster::triples() {
  CC4S_ASSERT(inputs_set_, "inputs not set");
  CC4S_ASSERT(have_amplitudes_, "no amplitudes yet");
  CC4S_HIP(hipSetDevice(device_));
  return ccsd_->triples(T_.get(), static_cast<const double *>(eps_h_.get()),
                        static_cast<const double *>(eps_p_.get()));
}

// --- energy adjoint ----------------------------------------------------------

const void *CoupledCluster::adjoint_vertex(const double *gamma, bool on_device) {
  if (!gamma) return gamma_.get();
  if (on_device) return gamma;
  const auto &s = p_.shape;
  const size_t bytes = static_cast<size_t>(s.NG) * s.Np() * s.Np() * 16;
  adj_gamma_.reserve(bytes);
  CC4S_HIP(hipMemcpyAsync(adj_gamma_.get(), gamma, bytes, hipMemcpyHostToDevice, stream_));
  return adj_gamma_.get();
}

// hp block [G, i, a] or ph block [G, b, j] of a vertex; the real field reads
// it as real numbers over the (Re, Im) pairs, [2 NG, ., .]
View CoupledCluster::vertex_block(const void *gamma, bool hp,
                                  const std::string &names) const {
  const auto &s = p_.shape;
  const int64_t No = s.No, Np = s.Np(), NG = s.NG;
  View full;
  if (s.field == Field::Complex) {
    full = dense_view(const_cast<void *>(gamma), Field::Complex, "GPQ", {NG, Np, Np});
  } else {
    full.ptr = const_cast<void *>(gamma);
    full.field = Field::Real;
    full.idx = "GPQ";
    full.lens = {2 * NG, Np, Np};
    full.strides = {1, 2 * NG, 2 * NG * Np};
  }
  const View v = hp ? full.sliced(1, 0, No).sliced(2, No, Np)
                    : full.sliced(1, No, Np).sliced(2, 0, No);
  return v.as(names);
}

// S[G] = Re sum_ai conj(Gamma^G_ai) Y^G_ai / V[G] with Y^G_ai = sum_bj
// Gamma^G_jb Z_abij, Z = 2 tau_abij - tau_baij: the reference's two terms
// 2 cTC_Gia C_Gjb tau_abij - cTC_Gja C_Gib tau_abij with i <-> j renamed in
// the second (tau_abji = tau_baij), C = Gamma / sqrt(V).
void CoupledCluster::transition_structure_factor(const double *U, int64_t NGg,
                                                 const double *V, double *S) {
  CC4S_ASSERT(inputs_set_, "inputs not set");
  CC4S_ASSERT(have_amplitudes_, "no amplitudes yet");
  CC4S_ASSERT(p_.shape.field == Field::Complex,
              "the transition structure factor needs the complex field");
  CC4S_ASSERT(V && S, "null argument");
  CC4S_HIP(hipSetDevice(device_));
  const auto &s = p_.shape;
  const int64_t No = s.No, Nv = s.Nv, NG = s.NG;
  CC4S_ASSERT(U ? NGg > 0 : NGg == NG,
              "without singular vectors the grid has the vertex's NG points");
  const size_t block = static_cast<size_t>(NGg) * No * Nv * 16;
  energy_weights(stream_, s.field, T_.get(), No, Nv, R_.get());
  const View Z = dense_view(R_.get(), s.field, "abij", {Nv, Nv, No, No});
  View ph, hp;
  if (U) {
    fsc_u_.reserve(static_cast<size_t>(NGg) * NG * 16);
    CC4S_HIP(hipMemcpyAsync(fsc_u_.get(), U, static_cast<size_t>(NGg) * NG * 16,
                            hipMemcpyHostToDevice, stream_));
    fsc_ph_.reserve(block);
    fsc_hp_.reserve(block);
    const View Uv = dense_view(fsc_u_.get(), Field::Complex, "GF", {NGg, NG});
    ph = dense_view(fsc_ph_.get(), Field::Complex, "Gai", {NGg, Nv, No});
    hp = dense_view(fsc_hp_.get(), Field::Complex, "Gjb", {NGg, No, Nv});
    engine_->contract(1.0, Uv, vertex_block(gamma_.get(), false, "Fai"), 0.0, ph);  // :71-80
    engine_->contract(1.0, Uv, vertex_block(gamma_.get(), true, "Fjb"), 0.0, hp);
  } else {
    ph = vertex_block(gamma_.get(), false, "Gai");
    hp = vertex_block(gamma_.get(), true, "Gjb");
  }
  fsc_y_.reserve(block);
  const View Y = dense_view(fsc_y_.get(), Field::Complex, "Gai", {NGg, Nv, No});
  engine_->contract(1.0, hp, Z, 0.0, Y);
  adj_partials_.reserve(block_dots_partials_size(NGg));
  adj_S_.reserve(NGg * sizeof(double));
  if (U)  // dense [NGg, (a,i)] blocks
    block_dots(stream_, fsc_ph_.get(), NGg, Nv * No, 0, Nv * No, 0, 1, fsc_y_.get(), 1.0,
               false, adj_partials_.get(), adj_S_.as<double>());
  else
    block_dots(stream_, gamma_.get(), NG, s.Np(), No, Nv, 0, No, fsc_y_.get(), 1.0, false,
               adj_partials_.get(), adj_S_.as<double>());
  CC4S_HIP(hipMemcpyAsync(S, adj_S_.get(), NGg * sizeof(double), hipMemcpyDeviceToHost,
                          stream_));
  CC4S_HIP(hipStreamSynchronize(stream_));
  for (int64_t G = 0; G < NGg; ++G) S[G] /= V[G];  // (1 / sqrt V)^2, :87-102
}

// E = sum_G Re sum_ia conj(Gamma^G_ia) g^G_ia with g^G_ia = sum_bj conj(Z_abij)
// Gamma^G_bj (complex), or 1/2 sum_G sum_ai Re(conj Gamma^G_ai g^G_ai) with
// g^G_ai = sum_bj Zs_abij Gamma^G_bj (real): the pphh integrals are never
// formed, and Z exists only tile by tile inside the product kernel, rebuilt
// from the amplitudes.
double CoupledCluster::energy_forward(const double *gamma, bool gamma_on_device) {
  CC4S_ASSERT(inputs_set_, "inputs not set");
  CC4S_ASSERT(have_amplitudes_, "no amplitudes yet");
  CC4S_HIP(hipSetDevice(device_));
  const auto &s = p_.shape;
  const int64_t No = s.No, Nv = s.Nv, NG = s.NG;
  const void *g = adjoint_vertex(gamma, gamma_on_device);
  const size_t block_bytes = static_cast<size_t>(NG) * No * Nv * 16;
  adj_partials_.reserve(block_dots_partials_size(NG));
  adj_S_.reserve(NG * sizeof(double));
  if (s.field == Field::Complex) {
    adj_hp_.reserve(block_bytes);
    energy_gradient_product(stream_, s.field, false, GradientBlock::HP, g, T_.get(), NG,
                            No, Nv, adj_hp_.get());
    block_dots(stream_, g, NG, s.Np(), 0, No, No, Nv, adj_hp_.get(), 1.0, false,
               adj_partials_.get(), adj_S_.as<double>());
  } else {
    adj_ph_.reserve(block_bytes);
    energy_gradient_product(stream_, s.field, false, GradientBlock::PH, g, T_.get(), NG,
                            No, Nv, adj_ph_.get());
    block_dots(stream_, g, NG, s.Np(), No, Nv, 0, No, adj_ph_.get(), 0.5, false,
               adj_partials_.get(), adj_S_.as<double>());
  }
  std::vector<double> S(NG);
  CC4S_HIP(hipMemcpyAsync(S.data(), adj_S_.get(), NG * sizeof(double),
                          hipMemcpyDeviceToHost, stream_));
  CC4S_HIP(hipStreamSynchronize(stream_));
  double E = 0.0;
  for (double v : S) E += v;
  return E;
}

// Complex field: g_hp^G_ia = sum_bj conj(Z_abij) Gamma^G_bj and
// g_ph^G_bj = sum_ai Z_abij Gamma^G_ia; real field: g_ph^G_ai =
// sum_bj Zs_abij Gamma^G_bj; every other block of the gradient is zero.
void CoupledCluster::energy_backward(const double *gamma, bool gamma_on_device,
                                     double *grad, bool grad_on_device, double *S,
                                     bool fp32) {
  CC4S_ASSERT(inputs_set_, "inputs not set");
  CC4S_ASSERT(have_amplitudes_, "no amplitudes yet");
  CC4S_HIP(hipSetDevice(device_));
  const auto &s = p_.shape;
  const int64_t No = s.No, Nv = s.Nv, NG = s.NG, Np = s.Np();
  const bool cplx_field = s.field == Field::Complex;
  const void *g = adjoint_vertex(gamma, gamma_on_device);
  const size_t block_bytes = static_cast<size_t>(NG) * No * Nv * 16;
  if (cplx_field) adj_hp_.reserve(block_bytes);
  adj_ph_.reserve(block_bytes);
  // the energy weights are rebuilt from the amplitudes tile by tile inside
  // the products (FP64 or FP32 MFMA)
  if (cplx_field)
    energy_gradient_product(stream_, s.field, fp32, GradientBlock::HP, g, T_.get(), NG, No,
                            Nv, adj_hp_.get());
  energy_gradient_product(stream_, s.field, fp32, GradientBlock::PH, g, T_.get(), NG, No, Nv,
                          adj_ph_.get());
  if (S) {
    adj_partials_.reserve(block_dots_partials_size(NG));
    adj_S_.reserve(NG * sizeof(double));
    double *Sd = adj_S_.as<double>();
    if (cplx_field) {
      block_dots(stream_, g, NG, Np, 0, No, No, Nv, adj_hp_.get(), 0.5, false,
                 adj_partials_.get(), Sd);
      block_dots(stream_, g, NG, Np, No, Nv, 0, No, adj_ph_.get(), 0.5, true,
                 adj_partials_.get(), Sd);
    } else {
      block_dots(stream_, g, NG, Np, No, Nv, 0, No, adj_ph_.get(), 0.5, false,
                 adj_partials_.get(), Sd);
    }
    CC4S_HIP(hipMemcpyAsync(S, Sd, NG * sizeof(double), hipMemcpyDeviceToHost, stream_));
  }
  if (grad) {
    // element (G, p, q) of the gradient at G + NG (p + Np q); the hp block is
    // Nv runs of NG*No elements, the ph block No runs of NG*Nv elements
    const size_t bytes = static_cast<size_t>(NG) * Np * Np * 16;
    const hipMemcpyKind kind = grad_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    if (grad_on_device) {
      CC4S_HIP(hipMemsetAsync(grad, 0, bytes, stream_));
    } else {
      CC4S_HIP(hipStreamSynchronize(stream_));
      std::memset(grad, 0, bytes);
    }
    const size_t pitch = static_cast<size_t>(NG) * Np * 16;
    if (cplx_field)
      CC4S_HIP(hipMemcpy2DAsync(grad + 2 * NG * Np * No, pitch, adj_hp_.get(),
                                static_cast<size_t>(NG) * No * 16,
                                static_cast<size_t>(NG) * No * 16, Nv, kind, stream_));
    CC4S_HIP(hipMemcpy2DAsync(grad + 2 * NG * No, pitch, adj_ph_.get(),
                              static_cast<size_t>(NG) * Nv * 16,
                              static_cast<size_t>(NG) * Nv * 16, No, kind, stream_));
  }
  CC4S_HIP(hipStreamSynchronize(stream_));
}

}  // namespace cc4s_amd
