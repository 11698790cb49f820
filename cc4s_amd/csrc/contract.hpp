// Tensor contraction engine: C[c] = alpha * A[a] B[b] + beta * C[c] with
// Einstein summation over the indices absent from C.  This is the semantics
// of one tcc contraction statement (tcc/Contraction.hpp:160-477, executed by
// engines/CtfMachineTensor.hpp:117-133 in the reference): "<<=" is beta = 0,
// "+=" is beta = 1, a leading scalar is alpha.
//
// Every contraction is executed by one hand-written gfx950 kernel family: a
// generalised tensor-tensor contraction (GETT) on FP64 MFMA
// (v_mfma_f64_16x16x4_f64), tiles staged through LDS, operands addressed
// through per-mode offset tables so that any index permutation is read in
// place (no transpose kernels), optional split-K with a deterministic
// reduction.  Complex operands use four real MFMA products per complex MAC.
#pragma once

#include "tensor.hpp"

#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

namespace cc4s_amd {

struct GettPlan;

struct ContractStats {
  double flops = 0;       // algorithmic flops (8 per complex MAC, 2 per real)
  int64_t launches = 0;
};

// per kernel variant: launches, summed duration (HIP events around the GETT
// launch on the engine's stream, split-K reduction excluded) and flops
struct KernelTiming {
  int64_t launches = 0;
  double ms = 0, flops = 0;
};

class ContractionEngine {
 public:
  explicit ContractionEngine(hipStream_t s);
  ~ContractionEngine();

  // C = alpha * op(A) op(B) + beta * C, op = conjugation where the view says
  void contract(double alpha, const View &A, const View &B, double beta,
                const View &C);
  // C = alpha * op(A1) op(B1) + alpha2 * op(A2) op(B2) + beta * C for two
  // contractions whose operands split C's indices the same way: one launch
  // with the second contraction as a second K segment, so C is read and
  // written once (complex FP64 passes; otherwise two launches)
  void contract2(double alpha, const View &A1, const View &B1, double alpha2,
                 const View &A2, const View &B2, double beta, const View &C);
  // C = alpha * op(A) + beta * C, A's indices are a permutation of C's
  void sum(double alpha, const View &A, double beta, const View &C);
  // C = beta * C
  void scale(double beta, const View &C);

  hipStream_t stream() const { return stream_; }
  ContractStats stats;

  // per-launch timing of the GETT kernels (off by default)
  void set_profiling(bool on);
  // resolves pending events; returns timings keyed by kernel variant name
  const std::unordered_map<std::string, KernelTiming> &timings();
  void reset_timings();
  // upper bound for split-K workspace (bytes)
  size_t max_workspace = size_t(1) << 30;

  // Mixed precision: complex contractions of at least `mixed_min_flops`
  // algorithmic flops run on the split-bf16 MFMA kernel (FP32 accumulation
  // per K-chunk, FP64 result); everything else stays FP64.
  void set_mixed(bool on) { mixed_ = on; }
  bool mixed() const { return mixed_; }
  double mixed_min_flops = 1e9;
  size_t pack_budget = size_t(1) << 30;  // bytes of packed operands per chunk
  // longest FP32 sum of the split-bf16 kernel (K is chunked beyond it and
  // the chunk results are added in FP64)
  int64_t mixed_max_k = 16384;
  // split parts in IEEE half with per-operand power-of-two scaling (about
  // 22 significant bits per real part) instead of bf16 (about 16)
  bool mixed_half = false;
  // 16-bit products per complex multiply-add of the split kernel: 12 (the
  // hi*hi, hi*lo and lo*hi products of both real parts) or 16 (also lo*lo,
  // which lies 2^-22 below the product, under the FP32 sums' rounding), or,
  // with half parts, 9 (three real products per complex multiply-add: Re Re,
  // Im Im and (Re + Im)(Re + Im), each split in three)
  int mixed_terms = 12;
  // CC4S_MIXED_SPLIT, CC4S_MIXED_MAX_K, CC4S_MIXED_TERMS and
  // CC4S_GETT_GENERATION, when set, replace the settings above and the
  // complex kernel generation
  void apply_env_overrides();
  // While set, the split-bf16 packing of an operand whose memory starts at
  // `ptr` is kept for the whole K range and reused by later contractions of
  // the same plan; the caller guarantees that the memory does not change
  // until the slot is set again or cleared.  A non-empty `idx` also requires
  // the operand view's index names (two views of the same memory in one
  // contraction).  Three slots: the ladder keeps tau (its contractions'
  // constant operand; in the paired ladder the symmetric combination) in
  // slot 0, the right-hand dressed-vertex slice of a slice pair row (its
  // formations' constant operand) in slot 1 and the antisymmetric tau
  // combination of the paired ladder in slot 2.
  static constexpr int kCacheSlots = 3;
  void set_operand_cache(const void *ptr, int slot = 0, const std::string &idx = "") {
    cache_ptr_[slot] = ptr;
    cache_idx_[slot] = idx;
    cache_plan_[slot] = nullptr;
  }
  void clear_operand_cache() {
    for (int s = 0; s < kCacheSlots; ++s) {
      cache_ptr_[s] = nullptr;
      cache_idx_[s].clear();
      cache_plan_[s] = nullptr;
    }
  }

 private:
  GettPlan &plan_for(const View &A, const View &B, const View &C);
  GettPlan &plan_for2(const View &A1, const View &B1, const View &A2, const View &B2,
                      const View &C);
  void contract_mixed(GettPlan &P, double alpha, const View &A, const View &B,
                      double beta, const View &C, double flops);
  bool mixed_ = false;
  DeviceBuffer pack_;
  DeviceBuffer cache_[kCacheSlots];
  DeviceBuffer scales_;  // half split: [A, B, cache slots 0-2] scales, partials
  const void *cache_ptr_[kCacheSlots] = {};
  std::string cache_idx_[kCacheSlots];
  const GettPlan *cache_plan_[kCacheSlots] = {};
  bool cache_is_a_[kCacheSlots] = {true, true, true};
  size_t cache_eb_[kCacheSlots] = {};  // packed bytes per element and k
  struct Pending {
    hipEvent_t start, stop;
    std::string name;
    double flops;
  };
  bool profiling_ = false;
  std::vector<Pending> pending_;
  std::vector<hipEvent_t> event_pool_;
  hipEvent_t take_event();
  std::unordered_map<std::string, KernelTiming> timings_;
  hipStream_t stream_;
  std::unordered_map<std::string, std::unique_ptr<GettPlan>> plans_;
  DeviceBuffer workspace_;
};

// kernel-level entry used by the planner and by tests of the kernel itself
struct GettLaunch {
  Field field;
  const void *A, *B;
  void *C;
  void *W;  // split-K partials, splits * M * N elements, or null
  const int64_t *rowA, *colA, *rowB, *colB, *rowC, *colC;
  int64_t M, N, K;
  int splits;
  double alpha, beta;
  bool conjA, conjB;
  bool a_kcontig, b_kcontig;
  // optional second K segment (complex generation-3 kernel): k in
  // [K1p, K1p + K2) from A2, B2 and their tables, times f2; K = K1p + K2,
  // K1p = K1 rounded up to whole K-stages (16)
  const void *A2 = nullptr, *B2 = nullptr;
  const int64_t *rowA2 = nullptr, *colA2 = nullptr, *rowB2 = nullptr, *colB2 = nullptr;
  int64_t K1 = 0, K1p = 0, K2 = 0;
  double f2 = 1.0;
  bool conjA2 = false, conjB2 = false;
};
// `start`/`stop` (optional) bracket the main kernel, not the reduction
void launch_gett(const GettLaunch &L, hipStream_t s, hipEvent_t start = nullptr,
                 hipEvent_t stop = nullptr);
// name of the kernel variant launch_gett uses for L
std::string gett_variant(const GettLaunch &L);
// selects the complex kernel generation (1: 4 waves, 128x64 tile, single
// LDS buffer; 2: 8 waves, 128x128 tile, double-buffered LDS, four MFMA
// products per complex MAC; 3 (default): 8 waves of 32x32 on a 128x64 tile,
// three products per complex MAC; 4: as 3 with 4 waves of 64x32); plans
// created afterwards use its tile shape
void set_complex_gett_generation(int generation);

// split-K reduction shared by both kernel families:
// C[rowC[m] + colC[n]] = alpha * sum_s W[s][m][n] + beta * C
void launch_splitk_reduce(Field f, const double *W, double *C,
                          const int64_t *rowC, const int64_t *colC, int64_t M,
                          int64_t N, int splits, double alpha, double beta,
                          hipStream_t s);

// split-bf16 path (mixed.hip)
struct MixedTile {
  int BM, BN, BK;
};
// multiply tile of the split kernel: 2 (default) 128 x 128 with 4 waves, two
// workgroups per CU; 1: 128 x 256 with 8 waves, one workgroup per CU
// (CC4S_MIXED_TILE overrides); the 9-product kernel always uses 128 x 128
MixedTile mixed_tile(int terms = 12);
void set_mixed_tile(int tile);
// packs k in [k0, k0+Kc) of an operand (FP64 complex, addressed by
// offMN[mn] + offK[k]) into split words, zero-padded: bf16 parts, or
// (scale != null) half parts of the operand times scale[0].  Layout
// [Kc/2][MNp] of 16-byte pairs: pair (p, mn) holds the words of k = 2p and
// 2p+1, word = (Re hi | Re lo << 16, Im hi | Im lo << 16); Kc is even.
// out_s (half only, or null): [Kc/2][MNp] 8-byte pairs of the words of
// Re + Im (of the conjugate when conj) of k = 2p and 2p+1.
void launch_pack(const double *X, const int64_t *offMN, const int64_t *offK,
                 int64_t MN, int64_t K, int64_t k0, int64_t Kc, int64_t MNp,
                 bool conj, bool kcontig, void *out, void *out_s,
                 const double *scale, hipStream_t s);
// scale[0] = the power of two that brings max(|Re|, |Im|) of the operand
// (max(|Re| + |Im|) with sumabs) into [2^14, 2^15) (the half split's range);
// partials holds split_scale_partials_bytes()
void launch_split_scale(const double *X, const int64_t *offMN, const int64_t *offK,
                        int64_t MN, int64_t K, bool kcontig, bool sumabs,
                        double *partials, double *scale, hipStream_t s);
size_t split_scale_partials_bytes();
// C = alpha * sum_s W[s] / (scaleA[0] scaleB[0]) + beta * C over the FP32
// partials of launch_bgemm (scales null: bf16 parts, no scaling)
void launch_split_reduce(const void *W, double *C, const int64_t *rowC,
                         const int64_t *colC, int64_t M, int64_t N, int splits,
                         double alpha, double beta, const double *scaleA,
                         const double *scaleB, hipStream_t s);
// C (or split-K partials W) from packed operands ([Kc/2][Mp] and [Kc/2][Np]
// pairs); scaleA/scaleB (both or neither): half parts, the sums are divided
// by both; terms: 16-bit products per complex multiply-add (12 or 16, or 9:
// half parts only, with the Re + Im planes Asum/Bsum of both operands).
// Partials W [splits][M][N]: with 12 or 16 terms the unscaled FP32 sums
// (8 bytes per element, reduced by launch_split_reduce), with 9 the scaled
// FP64 values (16 bytes, reduced by launch_splitk_reduce).
void launch_bgemm(const void *Apack, const void *Bpack, double *C, void *W,
                  const int64_t *rowC, const int64_t *colC, int64_t M, int64_t N,
                  int64_t Mp, int64_t Np, int64_t Kc, int splits, double alpha,
                  double beta, const double *scaleA, const double *scaleB,
                  int terms, const void *Asum, const void *Bsum, hipStream_t s);

}  // namespace cc4s_amd
