/*
 * C ABI of the MI355X CCSD library: a drop-in for the CoupledCluster step
 * with method Ccsd and mixer DiisMixer of a cc4s operation file.
 *
 * Reference interfaces replaced (paths relative to the cc4s source tree):
 *   CoupledCluster::run            src/main/algorithms/CoupledCluster.cxx:36-192
 *   CoupledClusterMethod::getResiduum (Ccsd)
 *                                  src/main/algorithms/coupledcluster/CoupledClusterMethod.hpp:49-51,
 *                                  src/main/algorithms/coupledcluster/Ccsd.cxx:76-633
 *   Mixer::append/get/getResiduumNorm (DiisMixer)
 *                                  src/main/mixers/Mixer.hpp:49-69, src/main/mixers/DiisMixer.cxx:58-117
 *   VertexCoulombIntegrals::run    src/main/algorithms/VertexCoulombIntegrals.cxx:37-197
 *   DefineHolesAndParticles / SliceOperator
 *                                  src/main/algorithms/DefineHolesAndParticles.cxx:38-96,
 *                                  src/main/algorithms/SliceOperator.cxx:54-131
 *
 * Conventions
 *   - Tensors cross the boundary in the cc4s layout: dense, first index
 *     fastest (reference/binary_tensor_file_format.txt:86-95).  Complex values
 *     are interleaved (re, im) doubles.  Gamma is [NG, Np, Np] (the
 *     CoulombVertex tensor), T1/R1 are [Nv, No] ("ph"), T2/R2 are
 *     [Nv, Nv, No, No] ("pphh").  Real-field amplitudes are plain doubles.
 *   - Every call returns 0 on success and a nonzero status otherwise;
 *     cc4s_ccsd_last_error() gives the message (the reference throws an
 *     exception with the same text, Cc4s.cxx:356-389).
 *   - The caller owns host buffers (copied in and out); the library owns
 *     device memory.  One HIP stream per context; calls are blocking and not
 *     re-entrant on one context (the reference's execute() is synchronous).
 */
#ifndef CC4S_CCSD_H
#define CC4S_CCSD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cc4s_ccsd_context cc4s_ccsd_context;

/* Scalar field: the reference runs Ccsd<Real> when the vertex metaData has
 * halfGrid: 1 and Ccsd<Complex> otherwise (CoupledCluster.cxx:36-52,
 * VertexCoulombIntegrals.cxx:45-52). */
enum { CC4S_FIELD_REAL = 0, CC4S_FIELD_COMPLEX = 1 };

/* Arithmetic of the large contractions (ladder, ring, Gamma terms).  FP64
 * reproduces the reference.  MIXED (complex field only) runs them on 16-bit
 * MFMA with every FP64 operand split into hi + lo parts (IEEE half with a
 * power-of-two scale per operand, or bf16: mixed_split) and FP32 sums of at
 * most mixed_max_k terms, until the mixer's dR falls below mixed_switch (or
 * stops falling: two passes without halving its best value).  Then, with
 * mixed_refine (default), one FP64 anchor pass also evaluates the split
 * residuum at the same amplitudes; its defect d = R64 - Rsplit corrects the
 * split passes that follow until they meet the convergence test or stall,
 * and FP64 passes confirm (mixed_refine 3: a corrected pass that meets the
 * convergence test ends the run); without it FP64 passes follow directly and
 * mixed_restart decides whether the DIIS history restarts at the switch.
 * Otherwise only FP64 passes can declare convergence.  Denominators, energy, DIIS and
 * every small contraction stay FP64.  FP32 is accepted by the energy
 * gradient only. */
enum { CC4S_PRECISION_FP64 = 0, CC4S_PRECISION_MIXED = 1, CC4S_PRECISION_FP32 = 2 };

/* Coupled-cluster method (CoupledCluster.cxx:77-84, registered names "Ccsd"
 * Ccsd.cxx:25-32 and "Drccd" Drccd.cxx:25-28). */
enum { CC4S_METHOD_CCSD = 0, CC4S_METHOD_DRCCD = 1 };

/* Step arguments of CoupledCluster (CoupledCluster.cxx:59-117,260-262,
 * CoupledCluster.hpp:45-49, Ccsd.cxx:95,287-289, DiisMixer.cxx:33). */
typedef struct {
  int field;                      /* CC4S_FIELD_* */
  int64_t No, Nv, NG;             /* holes, particles, auxiliary field */
  int64_t max_iterations;         /* maxIterations (16) */
  double energy_convergence;      /* energyConvergence (1e-7) */
  double amplitudes_convergence;  /* amplitudesConvergence (1e-6) */
  double level_shift;             /* levelShift (0) */
  int64_t max_residua;            /* mixer.maxResidua (4), at most 16 */
  int ppl;                        /* ppl (1) */
  int64_t integrals_slice_size;   /* integralsSliceSize (0 = No) */
  int precision;                  /* CC4S_PRECISION_* */
  double mixed_switch;            /* MIXED: dR below which FP64 takes over (1e-5) */
  int device;                     /* HIP device ordinal */
  int method;                     /* CC4S_METHOD_* (Ccsd) */
  int linearized;                 /* Drccd: linearized (0), Drccd.cxx:53-55 */
  int adjacent_pairs_exchange;    /* Drccd: adjacentPairsExchange (0), Drccd.cxx:56-58 */
  int mixed_restart;              /* MIXED: DIIS history at the switch: 0 keep, 1 restart,
                                     2 restart when dR < 2e-6 or stalled (2) */
  int mixed_restart_keep;         /* MIXED: most recent pairs a restart keeps (0) */
  int mixed_refine;               /* MIXED: after the split passes: 0 FP64 passes; 1, 2 an
                                     FP64 anchor pass that also evaluates the split residuum
                                     at the same amplitudes (defect d = R64 - Rsplit, DIIS
                                     restarted), split passes with Rsplit + d until they stop
                                     gaining, then FP64 passes (2: DIIS restarted before
                                     them, 1: kept); 3 as 1, but a corrected pass
                                     meeting the convergence test ends the run (1) */
  int mixed_split;                /* MIXED: split parts 0 bf16, 1 IEEE half with a
                                     power-of-two scale per operand (1) */
  int64_t mixed_max_k;            /* MIXED: longest FP32 sum of a split contraction; longer
                                     K is cut into chunks added in FP64 (2048) */
  int ladder;                     /* particle-particle ladder of complex passes
                                     (Ccsd.cxx:579-600): CC4S_LADDER_SLAB (V slab through
                                     HBM, as the reference), CC4S_LADDER_FUSED (FP64 passes:
                                     V formed in LDS inside the contraction kernel) or
                                     CC4S_LADDER_PAIRED (default: the slab contraction over
                                     pairs c <= d, i <= j of tau's symmetric and
                                     antisymmetric combinations, half its multiply-adds;
                                     FP64 and split passes); same results to rounding.
                                     The real field always uses the slab. */
} cc4s_ccsd_params;

/* Values of cc4s_ccsd_params.ladder. */
enum { CC4S_LADDER_SLAB = 0, CC4S_LADDER_FUSED = 1, CC4S_LADDER_PAIRED = 2 };

/* One pass of the iteration loop (CoupledCluster.cxx:129-175). */
typedef struct {
  int64_t iteration;       /* 1-based, the reference's "Iter" column */
  double energy;           /* Re E after mixing */
  double delta_energy;     /* Re (E - E_previous) */
  double residuum_norm;    /* dR of the mixer (DiisMixer.cxx:106) */
  double direct, exchange; /* Re D, Re X */
  double seconds;          /* wall time of the pass */
  double flops;            /* algorithmic flops of the pass (8 per complex MAC) */
  int converged;           /* |dE| < energyConvergence and dR < amplitudesConvergence */
  int precision_used;      /* CC4S_PRECISION_* */
  int diis_placement;      /* DIIS history: 0 in HBM, 1 amplitude slots in pinned
                              host memory, 2 every slot in pinned host memory */
  double residuum_seconds; /* wall time of the residuum */
  double mixer_seconds;    /* wall time of the amplitude update and DIIS step */
  int pass_kind;           /* 0 FP64, 1 split, 2 FP64 anchor (also evaluates the
                              split residuum), 3 split corrected by the anchor's defect */
} cc4s_ccsd_iteration;

/* Result map of CoupledCluster::run (CoupledCluster.cxx:184-191,245-248). */
typedef struct {
  double correlation, direct, exchange; /* energy.{correlation,direct,exchange} */
  double second_order;                  /* energy.secondOrder (NaN if absent) */
  double unit;                          /* energy.unit */
  int convergence_reached;              /* convergenceReached */
  int64_t iterations;                   /* passes done */
} cc4s_ccsd_result;

/* Reference defaults: CoupledCluster.hpp:45-49, DiisMixer.cxx:33, Ccsd.cxx:95. */
void cc4s_ccsd_default_params(cc4s_ccsd_params *params);

/* Creates a context (stream, contraction plans) on params->device. */
int cc4s_ccsd_create(const cc4s_ccsd_params *params, cc4s_ccsd_context **ctx);

/* Inputs of the step: the CoulombVertex Gamma [NG, Np, Np] complex
 * interleaved, and the eigen energies eps [Np] with holes first.  Replaces the
 * DefineHolesAndParticles, SliceOperator and VertexCoulombIntegrals steps
 * (the V slices are formed on the GPU here, once per run). */
int cc4s_ccsd_set_inputs(cc4s_ccsd_context *ctx, const double *gamma,
                         const double *eps);

/* Same, with Gamma already resident in device memory (copied device to
 * device; the caller keeps ownership of gamma_device). */
int cc4s_ccsd_set_inputs_device(cc4s_ccsd_context *ctx, const void *gamma_device,
                                const double *eps);

/* initialAmplitudes (CoupledCluster.cxx:70-75); T1 [Nv,No], T2 [Nv,Nv,No,No]. */
int cc4s_ccsd_set_amplitudes(cc4s_ccsd_context *ctx, const double *T1,
                             const double *T2);

/* One pass of the loop: residuum, denominator update, DIIS, energy. */
int cc4s_ccsd_iterate(cc4s_ccsd_context *ctx, cc4s_ccsd_iteration *out);

/* The whole loop until convergence or maxIterations, then the final energy
 * (CoupledCluster.cxx:129-191).  `trace` receives up to `trace_capacity`
 * passes (may be NULL). */
int cc4s_ccsd_run(cc4s_ccsd_context *ctx, cc4s_ccsd_result *out,
                  cc4s_ccsd_iteration *trace, int64_t trace_capacity);

/* Current amplitudes (the mixer's output) in cc4s layout. */
int cc4s_ccsd_get_amplitudes(cc4s_ccsd_context *ctx, double *T1, double *T2);

/* Ccsd::getResiduum at the given amplitudes (Ccsd.cxx:76-633); T1 == NULL
 * and T2 == NULL evaluate the first-iteration residuum (R1 = 0, R2 = V). */
int cc4s_ccsd_residuum(cc4s_ccsd_context *ctx, const double *T1,
                       const double *T2, double *R1, double *R2);

/* CoupledCluster::getEnergy of the given amplitudes (CoupledCluster.cxx:
 * 194-253): out[0..5] = Re E, Im E, Re D, Im D, Re X, Im X. */
int cc4s_ccsd_energy(cc4s_ccsd_context *ctx, const double *T1,
                     const double *T2, double *out);

/* FiniteSizeCorrection::calculateTransitionStructureFactor
 * (FiniteSizeCorrection.cxx:49-133) at the context's current amplitudes,
 * complex field: the vertex on the momentum grid Gamma^G = sum_F Gamma^F
 * U[G,F] (singular_vectors: U [NG_grid, NG] complex interleaved, first index
 * fastest, host memory; NULL when the vertex is on the grid, NG_grid = NG),
 * codensities C = Gamma^G / sqrt(V[G]) (coulomb_potential: V [NG_grid], host)
 * and structure_factor[G] = Re(2 conj(C_Gai) C_Gjb tau_abij -
 * conj(C_Gaj) C_Gib tau_abij), tau = T2 + T1 T1 (host, NG_grid values). */
int cc4s_ccsd_transition_structure_factor(cc4s_ccsd_context *ctx,
                                          const double *singular_vectors,
                                          int64_t NG_grid,
                                          const double *coulomb_potential,
                                          double *structure_factor);

/* FiniteSizeCorrection::interpolation (FiniteSizeCorrection.cxx:136-311),
 * host computation, no context: grid_vectors [NG][3] Cartesian momenta,
 * reciprocal_lattice [3][3] (rows: metaData Gi, Gj, Gk of the grid),
 * structure_factor [NG], N the fine-grid resolution
 * (interpolationGridSize, 20), and the units of the vertex, the grid and the
 * potential.  out[0] = corrected, out[1] = uncorrected; the correction is
 * out[0] - out[1]. */
int cc4s_finite_size_correction(const double *grid_vectors, int64_t NG,
                                const double *reciprocal_lattice,
                                const double *structure_factor, int64_t N,
                                double vertex_unit, double grid_unit,
                                double potential_unit, double *out);

/* (T) correlation energy of the context's current amplitudes (after run /
 * iterate, or set_amplitudes): PerturbativeTriplesReference::run
 * (PerturbativeTriplesReference.cxx:55-180, cT off), with the V_ppph slice
 * formed for the call.  *energy = Re E. */
int cc4s_ccsd_triples(cc4s_ccsd_context *ctx, double *energy);

/* Energy adjoint at the context's current amplitudes (after iterate/run or
 * set_amplitudes).  The reference has no such entry point: its finite-size
 * analysis (FiniteSizeCorrection.cxx:50-133) contracts the converged
 * amplitudes with the vertex into a transition structure factor; this pair
 * exposes the correlation energy as a function of the vertex and its exact
 * gradient, whose per-G split is that structure factor without the Coulomb
 * potential and singular-vector factors (FiniteSizeCorrection.cxx:74-124).
 *
 * forward: *energy = Re E of CoupledCluster.cxx:194-253 with the pphh
 * integrals formed from `gamma` as VertexCoulombIntegrals.cxx:68-171 forms
 * them.  gamma: [NG, Np, Np] complex interleaved, in host memory
 * (gamma_on_device 0) or device memory (1); NULL: the vertex of set_inputs. */
int cc4s_ccsd_energy_grad_forward(cc4s_ccsd_context *ctx, const double *gamma,
                                  int gamma_on_device, double *energy);
/* backward: grad[G,p,q] = dE/dRe Gamma^G_pq + i dE/dIm Gamma^G_pq ([NG, Np,
 * Np] complex interleaved, host or device memory per grad_on_device; NULL to
 * skip) and structure_factor[G] = 1/2 sum_pq Re(conj(Gamma^G_pq) grad^G_pq)
 * (host, NG values; NULL to skip), whose sum over G is the energy.
 * precision: CC4S_PRECISION_FP64 (FP64 MFMA) or CC4S_PRECISION_FP32 (FP32
 * MFMA operands and sums).  Both directions rebuild the energy weights
 * 2 tau_abij - tau_baij from the amplitudes tile by tile inside the gradient
 * products; no amplitude-sized intermediate is stored. */
int cc4s_ccsd_energy_grad_backward(cc4s_ccsd_context *ctx, const double *gamma,
                                   int gamma_on_device, double *grad,
                                   int grad_on_device, double *structure_factor,
                                   int precision);

/* Profiling (no counterpart in the reference, whose Timer/OperationsCounter
 * report only per-step totals, Cc4s.cxx:160-173): when enabled, every
 * contraction-kernel launch is bracketed by HIP events on the context's
 * stream, and the particle-particle ladder section is timed as a whole.
 * Enabling resets the counters. */
int cc4s_ccsd_set_profiling(cc4s_ccsd_context *ctx, int enabled);
int cc4s_ccsd_ladder_stats(cc4s_ccsd_context *ctx, double *milliseconds,
                           double *flops, int64_t *calls);
/* Launches, summed duration and algorithmic flops of one contraction-kernel
 * variant ("gett_c64_ak_bk", "gett_f64_am_bn", ...; a = A read along k (k)
 * or m (m), b = B read along k (k) or n (n)). */
int cc4s_ccsd_kernel_stats(cc4s_ccsd_context *ctx, const char *variant,
                           int64_t *launches, double *milliseconds,
                           double *flops);

/* Allocation-free energy.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  It computes what
 * cc4s_ccsd_energy computes (CoupledCluster::getEnergy,
 * CoupledCluster.cxx:194-253), with the same kernels and reduction order,
 * from device-resident buffers: T_device is one amplitude set [T1 (Nv*No);
 * T2 (Nv*Nv*No*No)] and Vpphh_device the V_abij slice [Nv,Nv,No,No] (complex
 * interleaved for CC4S_FIELD_COMPLEX), both in the cc4s layout, on the current
 * HIP device; stream is a hipStream_t (NULL: the default stream).  out[0..5] =
 * Re E, Im E, Re D, Im D, Re X, Im X (host memory).  No hipMalloc/hipFree:
 * the reduction partials live in a static 64 KiB device array per GPU (the
 * energy needs 32,800 bytes of it for any No, Nv), kept until the process
 * ends.  Calls that overlap on one GPU race on that array.  Errors:
 * cc4s_ccsd_last_error(NULL). */
int cc4s_ccsd_energy_noalloc(int field, int64_t No, int64_t Nv,
                             const void *T_device, const void *Vpphh_device,
                             void *stream, double *out);

/* Message of the last failed call on ctx (ctx == NULL: of the last failed
 * cc4s_ccsd_create, cc4s_contract or cc4s_ccsd_energy_noalloc). */
const char *cc4s_ccsd_last_error(const cc4s_ccsd_context *ctx);

void cc4s_ccsd_destroy(cc4s_ccsd_context *ctx);

/* One tensor contraction C[c] = alpha A[a] B[b] + beta C[c] on the GPU with
 * host buffers, Einstein summation over indices absent from c: the semantics
 * of one tcc statement (src/main/tcc/Contraction.hpp:160-477), exposed for
 * testing the contraction kernels.  lens_* give each index's length in the
 * order of the index string; field is CC4S_FIELD_*. */
int cc4s_contract(int field, double alpha, const double *A, const char *idx_a,
                  const int64_t *lens_a, int conj_a, const double *B,
                  const char *idx_b, const int64_t *lens_b, int conj_b,
                  double beta, double *C, const char *idx_c,
                  const int64_t *lens_c, int precision, int device);

#ifdef __cplusplus
}
#endif

#endif /* CC4S_CCSD_H */
