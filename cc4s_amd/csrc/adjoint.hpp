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

// Z in FP32 as the matrix M[k][n] = Z of the FP32 gradient products:
// complex field k = i + No*a, n = b + Nv*j, M = Z_abij;
// real field    k = a + Nv*i, n = b + Nv*j, M = Z_abij + Z_baji.
void energy_weights_fp32(hipStream_t s, Field f, const void *T, int64_t No,
                         int64_t Nv, void *M);

// FP32 copy of one block of the vertex: out[G + NG*(x + rows*y)] =
// Gamma[G, row0 + x, col0 + y] (complex, Gamma is [NG, Np, Np])
void vertex_block_fp32(hipStream_t s, const void *gamma, int64_t NG, int64_t Np,
                       int64_t row0, int64_t rows, int64_t col0, int64_t cols,
                       void *out);

// FP32 complex product C[G + NG*n] = sum_k A[G + NG*k] op(B)[k, n] with
// op(B)[k, n] = B[k*sbk + n*sbn] (conjugated if conjB); A, B complex FP32,
// C complex FP64; FP32 accumulation.
void gemm_fp32(hipStream_t s, const void *A, const void *B, int64_t sbk,
               int64_t sbn, bool conjB, void *C, int64_t NG, int64_t N,
               int64_t K);

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
