// (T) correction: the energy kernel (triples.hip); the contractions that
// form the connected triples amplitudes run on the contraction engine
// (Ccsd::triples).
#pragma once

#include "tensor.hpp"

namespace cc4s_amd {

// largest grid of the energy kernel (one running sum per block in `acc`);
// the complex kernel uses 768 blocks (three per CU, bound by LDS), the real
// one 1024 (four per CU, bound by threads): one resident wave of blocks, no
// tail
constexpr int kTriplesBlocks = 1024;

// acc[block] += partial sums of the weighted Re(W Zs)/D over W[a,b,c,k - k0],
// k in [k0, k0 + kc), for the hole pair i <= j <= k0; Vpphh [v,v,o,o],
// T1 [v,o]
void triples_energy(hipStream_t s, Field f, const void *W, const void *Vpphh,
                    const void *T1, const double *eps_h, const double *eps_p,
                    int64_t No, int64_t Nv, int64_t i, int64_t j, int64_t k0,
                    int64_t kc, double *acc);

}  // namespace cc4s_amd
