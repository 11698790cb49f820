// Bandwidth-bound kernels of the CCSD iteration (HBM-bound on MI355X).
#pragma once

#include "tensor.hpp"

#include <complex>
#include <vector>

namespace cc4s_amd {

using cplx = std::complex<double>;

// X_abij = T2_abij + f * T1_ai T1_bj   (Ccsd.cxx:460-463, 169-170, 215)
void tau(hipStream_t s, Field f, void *X, const void *T2, const void *T1,
         double factor, int64_t No, int64_t Nv);

// R_abij += R_baji with the old values on both sides (Ccsd.cxx:256, 520)
void symmetrize_pairs(hipStream_t s, Field f, void *R2, int64_t No, int64_t Nv);

// New amplitudes and amplitude change of one amplitude set [T1; T2]
// (CoupledCluster.cxx:137-144, 255-317):
//   R' = R - shift*T (only when T is given and shift != 0)
//   Tnew = R' * (-1/(Delta + shift)),  dT = Tnew - T (dT = Tnew without T)
// Delta_ai = eps_a - eps_i, Delta_abij = ((eps_a - eps_i) + eps_b) - eps_j.
void amplitude_update(hipStream_t s, Field f, const void *R, const void *T,
                      const double *eps_h, const double *eps_p, double shift,
                      void *Tnew, void *dT, int64_t No, int64_t Nv);

// Energy of amplitudes [T1;T2] against V_abij (CoupledCluster.cxx:204-235):
// D = 2 sum tau conj V_abij, X = - sum tau conj V_baij, tau = T2 + T1 T1.
// Returns (D, X); `partials` is device scratch of energy_partials_size().
size_t energy_partials_size();
void energy(hipStream_t s, Field f, const void *T, const void *Vpphh,
            int64_t No, int64_t Nv, void *partials, cplx *D, cplx *X);

// <x_k | y> for k < count over n elements (TensorSet.hpp:209-224).
size_t dots_partials_size(int count);
void multi_dot(hipStream_t s, Field f, const void *const *x, int count,
               const void *y, int64_t n, void *partials, std::vector<cplx> &out);

// out = sum_k c_k A_k ; returns <r|r> with r = sum_k c_k R_k
// (DiisMixer.cxx:92-106), r itself is never stored.
double combine(hipStream_t s, Field f, const void *const *A,
               const void *const *R, const std::vector<double> &c,
               int64_t n, void *out, void *partials);

// --- pair-packed amplitude sets ---------------------------------------------
// A packed set stores T1 (Nv*No elements) followed by T2 for the pairs
// P = a + Nv*i <= Q = b + Nv*j only, at Q*(Q+1)/2 + P; T2_abij = T2_baji
// makes the other half redundant.  Used for the DIIS history.
int64_t packed_set_size(int64_t No, int64_t Nv);

// amplitude_update writing packed T' and dT (reads R, T in full layout)
void amplitude_update_packed(hipStream_t s, Field f, const void *R, const void *T,
                             const double *eps_h, const double *eps_p,
                             double shift, void *Tnew, void *dT, int64_t No,
                             int64_t Nv);
// <x_k|y> of packed sets, equal to the full-layout product
void multi_dot_packed(hipStream_t s, Field f, const void *const *x, int count,
                      const void *y, int64_t No, int64_t Nv, void *partials,
                      std::vector<cplx> &out);
// out (full layout) = sum_k c_k A_k (packed); returns <r|r>, r = sum_k c_k R_k
double combine_packed(hipStream_t s, Field f, const void *const *A,
                      const void *const *R, const std::vector<double> &c,
                      int64_t No, int64_t Nv, void *out, void *partials);

// packed = alpha * pack(full) + beta * packed (beta 0: packed is not read)
void pack_axpby(hipStream_t s, Field f, const void *full, double alpha, void *packed,
                double beta, int64_t No, int64_t Nv);
// out (full layout) (first ? = : +=) c * x for the packed elements
// [t0, t0 + len) of one set, x = chunk[0 .. len); writes both mirror
// positions of a pair
void unpack_accumulate(hipStream_t s, Field f, const void *chunk, int64_t t0,
                       int64_t len, double c, bool first, void *out, int64_t No,
                       int64_t Nv);
// |sum_k c_k R_k|^2 of packed sets (the full-layout norm)
double norm_packed(hipStream_t s, Field f, const void *const *R,
                   const std::vector<double> &c, int64_t No, int64_t Nv,
                   void *partials);
// the two reductions above restricted to the packed elements [t0, t0 + len)
// of a set: x[k], y and R[k] point at element t0 (chunks of slots streamed
// from host memory); summing the chunks gives the whole-set value
void multi_dot_packed_range(hipStream_t s, Field f, const void *const *x, int count,
                            const void *y, int64_t t0, int64_t len, int64_t No,
                            int64_t Nv, void *partials, std::vector<cplx> &out);
double norm_packed_range(hipStream_t s, Field f, const void *const *R,
                         const std::vector<double> &c, int64_t t0, int64_t len,
                         int64_t No, int64_t Nv, void *partials);

// true when the complex vertex [NG, Np, Np] satisfies
// Gamma[G,q,p] == conj(Gamma[G,p,q]) exactly for every G, p, q
bool is_hermitian(hipStream_t s, const void *gamma, int64_t NG, int64_t Np);

// y = x (device copy on the stream)
void copy(hipStream_t s, void *y, const void *x, size_t bytes);
// y = a x + b y over n doubles
void axpby(hipStream_t s, int64_t n, double a, const void *x, double b, void *y);

}  // namespace cc4s_amd
