// extern "C" boundary (declared in include/cc4s_ccsd.h).
#include "../../include/cc4s_ccsd.h"

#include "coupled_cluster.hpp"
#include "finite_size.hpp"

#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <string>

using namespace cc4s_amd;

struct cc4s_ccsd_context {
  std::unique_ptr<CoupledCluster> cc;
  cc4s_ccsd_params params;
  std::string error;
};

namespace {

thread_local std::string g_error;

template <typename F>
int guarded(cc4s_ccsd_context *ctx, F &&body) {
  try {
    body();
    return 0;
  } catch (const std::exception &e) {
    (ctx ? ctx->error : g_error) = e.what();
  } catch (...) {
    (ctx ? ctx->error : g_error) = "unknown error";
  }
  return 1;
}

void fill(cc4s_ccsd_iteration *out, const IterationReport &r) {
  out->iteration = r.iteration;
  out->energy = r.energy.real();
  out->delta_energy = r.delta_energy;
  out->residuum_norm = r.residuum_norm;
  out->direct = r.direct.real();
  out->exchange = r.exchange.real();
  out->seconds = r.seconds;
  out->flops = r.flops;
  out->converged = r.converged ? 1 : 0;
  out->precision_used = r.mixed ? CC4S_PRECISION_MIXED : CC4S_PRECISION_FP64;
  out->diis_placement = r.diis_placement;
  out->residuum_seconds = r.residuum_seconds;
  out->mixer_seconds = r.mixer_seconds;
  out->pass_kind = r.kind;
}

}  // namespace

void cc4s_amd::set_last_error(const std::string &message) { g_error = message; }

extern "C" {

void cc4s_ccsd_default_params(cc4s_ccsd_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->field = CC4S_FIELD_COMPLEX;
  p->max_iterations = 16;
  p->energy_convergence = 1e-7;
  p->amplitudes_convergence = 1e-6;
  p->level_shift = 0.0;
  p->max_residua = 4;
  p->ppl = 1;
  p->integrals_slice_size = 0;
  p->precision = CC4S_PRECISION_FP64;
  p->mixed_switch = 1e-5;
  p->device = 0;
  p->mixed_restart = 2;
  p->mixed_refine = 1;
  p->mixed_split = 1;
  p->mixed_max_k = 2048;
  p->ladder = CC4S_LADDER_PAIRED;
}

int cc4s_ccsd_create(const cc4s_ccsd_params *params, cc4s_ccsd_context **out) {
  return guarded(nullptr, [&] {
    CC4S_ASSERT(params && out, "null argument");
    *out = nullptr;
    CC4S_ASSERT(params->field == CC4S_FIELD_REAL || params->field == CC4S_FIELD_COMPLEX,
                "unsupported tensor type as 'operator'");
    CC4S_ASSERT(params->precision == CC4S_PRECISION_FP64 ||
                    params->precision == CC4S_PRECISION_MIXED,
                "unknown precision mode");
    CC4S_ASSERT(params->precision == CC4S_PRECISION_FP64 ||
                    params->field == CC4S_FIELD_COMPLEX,
                "mixed precision needs the complex field");
    CoupledClusterParams p;
    p.mixed = params->precision == CC4S_PRECISION_MIXED;
    p.mixed_switch = params->mixed_switch;
    CC4S_ASSERT(params->method == CC4S_METHOD_CCSD || params->method == CC4S_METHOD_DRCCD,
                "Unknown method");
    p.ccsd.method = params->method == CC4S_METHOD_DRCCD ? Method::Drccd : Method::Ccsd;
    p.ccsd.linearized = params->linearized != 0;
    p.ccsd.adjacent_pairs_exchange = params->adjacent_pairs_exchange != 0;
    CC4S_ASSERT(params->mixed_restart >= 0 && params->mixed_restart <= 2,
                "mixed_restart must be 0, 1 or 2");
    p.mixed_restart = params->mixed_restart;
    CC4S_ASSERT(params->mixed_restart_keep >= 0 && params->mixed_restart_keep < params->max_residua,
                "mixed_restart_keep must lie in [0, maxResidua)");
    p.mixed_restart_keep = params->mixed_restart_keep;
    CC4S_ASSERT(params->mixed_refine >= 0 && params->mixed_refine <= 3,
                "mixed_refine must be 0, 1, 2 or 3");
    p.mixed_refine = params->mixed_refine;
    CC4S_ASSERT(params->mixed_split == 0 || params->mixed_split == 1,
                "mixed_split must be 0 (bf16) or 1 (half)");
    p.mixed_half = params->mixed_split == 1;
    CC4S_ASSERT(params->mixed_max_k >= 16, "mixed_max_k must be at least 16");
    p.mixed_max_k = params->mixed_max_k;
    p.shape.field = params->field == CC4S_FIELD_REAL ? Field::Real : Field::Complex;
    p.shape.No = params->No;
    p.shape.Nv = params->Nv;
    p.shape.NG = params->NG;
    p.max_iterations = params->max_iterations;
    p.energy_convergence = params->energy_convergence;
    p.amplitudes_convergence = params->amplitudes_convergence;
    p.level_shift = params->level_shift;
    p.max_residua = params->max_residua;
    p.ccsd.ppl = params->ppl != 0;
    p.ccsd.integrals_slice_size = params->integrals_slice_size;
    CC4S_ASSERT(params->ladder == CC4S_LADDER_SLAB || params->ladder == CC4S_LADDER_FUSED ||
                    params->ladder == CC4S_LADDER_PAIRED,
                "ladder must be CC4S_LADDER_SLAB, CC4S_LADDER_FUSED or CC4S_LADDER_PAIRED");
    p.ccsd.ladder_fused = params->ladder == CC4S_LADDER_FUSED;
    p.ccsd.ladder_paired = params->ladder == CC4S_LADDER_PAIRED;
    auto ctx = std::make_unique<cc4s_ccsd_context>();
    ctx->params = *params;
    ctx->cc.reset(new CoupledCluster(p, params->device));
    *out = ctx.release();
  });
}

int cc4s_ccsd_set_inputs(cc4s_ccsd_context *ctx, const double *gamma,
                         const double *eps) {
  if (!ctx) return 1;
  return guarded(ctx, [&] {
    CC4S_ASSERT(gamma && eps, "null input");
    ctx->cc->set_inputs(gamma, false, eps);
  });
}

int cc4s_ccsd_set_inputs_device(cc4s_ccsd_context *ctx, const void *gamma_device,
                                const double *eps) {
  if (!ctx) return 1;
  return guarded(ctx, [&] {
    CC4S_ASSERT(gamma_device && eps, "null input");
    ctx->cc->set_inputs(static_cast<const double *>(gamma_device), true, eps);
  });
}

int cc4s_ccsd_set_amplitudes(cc4s_ccsd_context *ctx, const double *T1,
                             const double *T2) {
  if (!ctx) return 1;
  return guarded(ctx, [&] {
    CC4S_ASSERT(T1 && T2, "null amplitudes");
    ctx->cc->set_amplitudes(T1, T2);
  });
}

int cc4s_ccsd_iterate(cc4s_ccsd_context *ctx, cc4s_ccsd_iteration *out) {
  if (!ctx) return 1;
  return guarded(ctx, [&] {
    IterationReport r = ctx->cc->iterate();
    if (out) fill(out, r);
  });
}

int cc4s_ccsd_run(cc4s_ccsd_context *ctx, cc4s_ccsd_result *out,
                  cc4s_ccsd_iteration *trace, int64_t trace_capacity) {
  if (!ctx) return 1;
  return guarded(ctx, [&] {
    std::vector<IterationReport> reps;
    const bool conv = ctx->cc->run(&reps);
    if (trace)
      for (int64_t k = 0; k < trace_capacity && k < static_cast<int64_t>(reps.size()); ++k)
        fill(trace + k, reps[k]);
    if (out) {
      out->correlation = ctx->cc->final_energy.real();
      out->direct = ctx->cc->final_direct.real();
      out->exchange = ctx->cc->final_exchange.real();
      out->second_order = ctx->cc->second_order;
      out->unit = 1.0;
      out->convergence_reached = conv ? 1 : 0;
      out->iterations = static_cast<int64_t>(reps.size());
    }
  });
}

int cc4s_ccsd_get_amplitudes(cc4s_ccsd_context *ctx, double *T1, double *T2) {
  if (!ctx) return 1;
  return guarded(ctx, [&] { ctx->cc->get_amplitudes(T1, T2); });
}

int cc4s_ccsd_residuum(cc4s_ccsd_context *ctx, const double *T1, const double *T2,
                       double *R1, double *R2) {
  if (!ctx) return 1;
  return guarded(ctx, [&] {
    CC4S_ASSERT((T1 == nullptr) == (T2 == nullptr), "give both T1 and T2 or neither");
    CC4S_ASSERT(R1 && R2, "null residuum buffers");
    ctx->cc->residuum(T1, T2, R1, R2);
  });
}

int cc4s_ccsd_energy(cc4s_ccsd_context *ctx, const double *T1, const double *T2,
                     double *out) {
  if (!ctx) return 1;
  return guarded(ctx, [&] {
    CC4S_ASSERT(T1 && T2 && out, "null argument");
    cplx E, D, X;
    ctx->cc->energy_of(T1, T2, &E, &D, &X);
    out[0] = E.real(); out[1] = E.imag();
    out[2] = D.real(); out[3] = D.imag();
    out[4] = X.real(); out[5] = X.imag();
  });
}

int cc4s_ccsd_triples(cc4s_ccsd_context *ctx, double *energy) {
  if (!ctx) return 1;
  return guarded(ctx, [&] {
    CC4S_ASSERT(energy, "null output");
    *energy = ctx->cc->triples();
  });
}

int cc4s_ccsd_transition_structure_factor(cc4s_ccsd_context *ctx,
                                          const double *singular_vectors,
                                          int64_t NG_grid,
                                          const double *coulomb_potential,
                                          double *structure_factor) {
  if (!ctx) return 1;
  return guarded(ctx, [&] {
    ctx->cc->transition_structure_factor(singular_vectors, NG_grid, coulomb_potential,
                                         structure_factor);
  });
}

int cc4s_finite_size_correction(const double *grid_vectors, int64_t NG,
                                const double *reciprocal_lattice,
                                const double *structure_factor, int64_t N,
                                double vertex_unit, double grid_unit,
                                double potential_unit, double *out) {
  return guarded(nullptr, [&] {
    CC4S_ASSERT(grid_vectors && reciprocal_lattice && structure_factor && out,
                "null argument");
    const FiniteSizeEnergy e =
        finite_size_correction(grid_vectors, NG, reciprocal_lattice, structure_factor, N,
                               vertex_unit, grid_unit, potential_unit);
    out[0] = e.corrected;
    out[1] = e.uncorrected;
  });
}

int cc4s_ccsd_energy_grad_forward(cc4s_ccsd_context *ctx, const double *gamma,
                                  int gamma_on_device, double *energy) {
  if (!ctx) return 1;
  return guarded(ctx, [&] {
    CC4S_ASSERT(energy, "null output");
    *energy = ctx->cc->energy_forward(gamma, gamma_on_device != 0);
  });
}

int cc4s_ccsd_energy_grad_backward(cc4s_ccsd_context *ctx, const double *gamma,
                                   int gamma_on_device, double *grad,
                                   int grad_on_device, double *structure_factor,
                                   int precision) {
  if (!ctx) return 1;
  return guarded(ctx, [&] {
    CC4S_ASSERT(precision == CC4S_PRECISION_FP64 || precision == CC4S_PRECISION_FP32,
                "unknown precision for the energy gradient");
    ctx->cc->energy_backward(gamma, gamma_on_device != 0, grad, grad_on_device != 0,
                             structure_factor, precision == CC4S_PRECISION_FP32);
  });
}

int cc4s_ccsd_set_profiling(cc4s_ccsd_context *ctx, int enabled) {
  if (!ctx) return 1;
  return guarded(ctx, [&] {
    Ccsd *m = ctx->cc->method();
    CC4S_ASSERT(m, "inputs not set");
    m->profile = enabled != 0;
    m->timings = CcsdTimings();
    ctx->cc->engine().set_profiling(enabled != 0);
    ctx->cc->engine().reset_timings();
  });
}

int cc4s_ccsd_kernel_stats(cc4s_ccsd_context *ctx, const char *variant,
                           int64_t *launches, double *ms, double *flops) {
  if (!ctx) return 1;
  return guarded(ctx, [&] {
    CC4S_ASSERT(variant, "null variant");
    const auto &t = ctx->cc->engine().timings();
    auto it = t.find(variant);
    const KernelTiming k = it == t.end() ? KernelTiming() : it->second;
    if (launches) *launches = k.launches;
    if (ms) *ms = k.ms;
    if (flops) *flops = k.flops;
  });
}

int cc4s_ccsd_ladder_stats(cc4s_ccsd_context *ctx, double *ms, double *flops,
                           int64_t *calls) {
  if (!ctx) return 1;
  return guarded(ctx, [&] {
    Ccsd *m = ctx->cc->method();
    CC4S_ASSERT(m, "inputs not set");
    if (ms) *ms = m->timings.ladder_ms;
    if (flops) *flops = m->timings.ladder_flops;
    if (calls) *calls = m->timings.ladder_calls;
  });
}

const char *cc4s_ccsd_last_error(const cc4s_ccsd_context *ctx) {
  return ctx ? ctx->error.c_str() : g_error.c_str();
}

void cc4s_ccsd_destroy(cc4s_ccsd_context *ctx) { delete ctx; }

int cc4s_contract(int field, double alpha, const double *A, const char *idx_a,
                  const int64_t *lens_a, int conj_a, const double *B,
                  const char *idx_b, const int64_t *lens_b, int conj_b,
                  double beta, double *C, const char *idx_c,
                  const int64_t *lens_c, int precision, int device) {
  return guarded(nullptr, [&] {
    CC4S_ASSERT(A && B && C && idx_a && idx_b && idx_c, "null argument");
    const Field f = field == CC4S_FIELD_REAL ? Field::Real : Field::Complex;
    CC4S_HIP(hipSetDevice(device));
    auto make = [&](const double *host, const char *idx, const int64_t *lens,
                    Tensor &t) {
      std::vector<int64_t> l(lens, lens + std::strlen(idx));
      t.reset(f, idx, l);
      if (t.bytes())
        CC4S_HIP(hipMemcpy(t.buf.get(), host, t.bytes(), hipMemcpyHostToDevice));
    };
    Tensor tA, tB, tC;
    make(A, idx_a, lens_a, tA);
    make(B, idx_b, lens_b, tB);
    make(C, idx_c, lens_c, tC);
    hipStream_t s;
    CC4S_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    {
      ContractionEngine E(s);
      View a = tA.v, b = tB.v;
      a.conj = conj_a != 0;
      b.conj = conj_b != 0;
      CC4S_ASSERT(precision == CC4S_PRECISION_FP64 ||
                      (precision == CC4S_PRECISION_MIXED && f == Field::Complex),
                  "mixed precision needs the complex field");
      E.set_mixed(precision == CC4S_PRECISION_MIXED);
      E.mixed_min_flops = 0.0;
      E.contract(alpha, a, b, beta, tC.v);
      CC4S_HIP(hipStreamSynchronize(s));
    }
    (void)hipStreamDestroy(s);
    if (tC.bytes())
      CC4S_HIP(hipMemcpy(C, tC.buf.get(), tC.bytes(), hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
