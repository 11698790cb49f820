// Kernels of the energy adjoint: the correlation energy at fixed amplitudes
// as a function of the Coulomb vertex, its gradient, and the per-G split of
// the energy (transition structure factor).
//
// The energy of CoupledCluster.cxx:194-253 is E = sum_abij Z_abij conj V_abij
// with Z_abij = 2 tau_abij - tau_baij, tau = T2 + T1 T1, and V_abij the pphh
// Coulomb integrals of VertexCoulombIntegrals.cxx:
//   complex: V_abij = sum_G conj(Gamma^G_ia) Gamma^G_bj          (:143-171)
//   real:    V_abij = sum_G Re Gamma^G_ai Re Gamma^G_bj + Im Im   (:68-102)
// E depends on the hp and ph blocks of Gamma only; it is homogeneous of
// degree two in Gamma, so S(G) = 1/2 sum_pq Re(conj(Gamma^G_pq) g^G_pq) with
// g = dE/dRe Gamma + i dE/dIm Gamma sums to E over G.
#pragma once

#include "tensor.hpp"

namespace cc4s_amd {

// Z (FP64) in the layout of T2 [a,b,i,j] from the amplitude set T = [T1; T2]:
// complex field Z_abij = 2 tau_abij - tau_baij; real field the symmetrised
// weight Z_abij + Z_baji (the real vertex enters V twice through the same
// block).
void energy_weights(hipStream_t s, Field f, const void *T, int64_t No,
                    int64_t Nv, void *Z);

// One block of the energy gradient, with the weights recomputed from the
// amplitude set T = [T1; T2] tile by tile inside the product (no Z tensor):
//   complex HP: C[G, i, a] = sum_bj conj(Z_abij) Gamma^G_bj
//   complex PH: C[G, b, j] = sum_ai Z_abij Gamma^G_ia
//   real    PH: C[G, a, i] = sum_bj (Z_abij + Z_baji) Gamma^G_bj
// (complex FP64 output, G fastest; for the real field Re and Im of C are the
// gradient of the two real vertex components).  FP64 (fp32 false) or FP32
// MFMA arithmetic with FP32 accumulation.
enum class GradientBlock { HP, PH };
void energy_gradient_product(hipStream_t s, Field f, bool fp32, GradientBlock blk,
                             const void *gamma, const void *T, int64_t NG, int64_t No,
                             int64_t Nv, void *C);

// Per-G dots over one block: out[G] (=, or += when accumulate) factor *
// Re sum_{x<rows, y<cols} conj(Gamma[G, row0+x, col0+y]) g[G + NG*(x + rows*y)]
// (complex storage for both fields; the real field's Re/Im parts are its two
// real vertex components).  Deterministic; `partials` holds
// block_dots_partials_size(NG) bytes.
size_t block_dots_partials_size(int64_t NG);
void block_dots(hipStream_t s, const void *gamma, int64_t NG, int64_t Np,
                int64_t row0, int64_t rows, int64_t col0, int64_t cols,
                const void *g, double factor, bool accumulate, void *partials,
                double *out);

}  // namespace cc4s_amd
