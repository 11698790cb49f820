"""Second, independent CCSD residuum formulation from explicit V slices.

TEST INFRASTRUCTURE ONLY (imported by tests/ to pin the oracle).

Restates CcsdReference<F>::getResiduum (coupledcluster/CcsdReference.cxx:
87-239): the same closed-shell CCSD equations written with the full set of
Coulomb integral slices, including V_pppp, instead of the Gamma-factorised,
dressed-vertex form of Ccsd.cxx.  It shares nothing with ccsd_oracle.py's
residuum functions except the V-slice formation, so agreement of the two at
1e-12 pins every term of the factorised form (a dropped term, a wrong sign or
a transposed operand shows up as a mismatch).  Sizes must stay tiny: V_pppp is
materialised.
"""

from __future__ import annotations

import numpy as np

from oracle.ccsd_oracle import ein, vertex_integrals

SLICES = ("pppp", "phph", "hhhh", "hhhp", "hhpp", "pphp", "phhh", "hphp",
          "phhp", "phpp", "hhph", "hppp", "pphh")


def residuum_vbased(slices, field, amplitudes, No, Nv):
    V = vertex_integrals(slices, field, SLICES)
    Vpphh = V["pphh"]
    dtype = Vpphh.dtype
    if amplitudes is None:                                       # :59-64
        return np.zeros((Nv, No), dtype=dtype), Vpphh.copy()
    T1, T2 = amplitudes
    Vpppp, Vphph, Vhhhh, Vhhhp = V["pppp"], V["phph"], V["hhhh"], V["hhhp"]
    Vhhpp, Vpphp, Vphhh, Vhphp = V["hhpp"], V["pphp"], V["phhh"], V["hphp"]
    Vphhp, Vphpp, Vhhph, Vhppp = V["phhp"], V["phpp"], V["hhph"], V["hppp"]

    Kac = (-2.0 * ein("klcd,adkl->ac", Vhhpp, T2)
           + ein("kldc,adkl->ac", Vhhpp, T2)
           - 2.0 * ein("klcd,ak,dl->ac", Vhhpp, T1, T1)
           + ein("kldc,ak,dl->ac", Vhhpp, T1, T1))               # :122-125
    Lac = (Kac + 2.0 * ein("akcd,dk->ac", Vphpp, T1)
           - ein("akdc,dk->ac", Vphpp, T1))                      # :128-130
    Kki = (2.0 * ein("klcd,cdil->ki", Vhhpp, T2)
           - ein("kldc,cdil->ki", Vhhpp, T2)
           + 2.0 * ein("klcd,ci,dl->ki", Vhhpp, T1, T1)
           - ein("kldc,ci,dl->ki", Vhhpp, T1, T1))               # :133-136
    Lki = (Kki + 2.0 * ein("klic,cl->ki", Vhhhp, T1)
           - ein("klci,cl->ki", Vhhph, T1))                      # :139-141
    R2 = ein("ac,cbij->abij", Lac, T2)                           # :144
    R2 -= ein("ki,abkj->abij", Lki, T2)                          # :147
    R2 += ein("abic,cj->abij", Vpphp, T1)                        # :150
    R2 -= ein("kbic,ak,cj->abij", Vhphp, T1, T1)                 # :151
    R2 -= ein("akij,bk->abij", Vphhh, T1)                        # :152
    R2 -= ein("akic,cj,bk->abij", Vphhp, T1, T1)                 # :153
    Xakic = (Vphhp.copy()
             - ein("lkic,al->akic", Vhhhp, T1)
             + ein("akdc,di->akic", Vphpp, T1)
             - 0.5 * ein("lkdc,dail->akic", Vhhpp, T2)
             - ein("lkdc,di,al->akic", Vhhpp, T1, T1)
             + ein("lkdc,adil->akic", Vhhpp, T2)
             - 0.5 * ein("lkcd,adil->akic", Vhhpp, T2))          # :156-162
    R2 += 2.0 * ein("akic,cbkj->abij", Xakic, T2)                # :163
    R2 -= ein("akic,bckj->abij", Xakic, T2)                      # :164
    Xakci = (Vphph.copy()
             - ein("lkci,al->akci", Vhhph, T1)
             + ein("akcd,di->akci", Vphpp, T1)
             - 0.5 * ein("lkcd,dail->akci", Vhhpp, T2)
             - ein("lkcd,di,al->akci", Vhhpp, T1, T1))           # :167-171
    R2 -= ein("akci,cbkj->abij", Xakci, T2)                      # :172
    R2 -= ein("bkci,ackj->abij", Xakci, T2)                      # :173
    R2 = R2 + R2.transpose(1, 0, 3, 2)                           # :176
    R2 += Vpphh                                                  # :184
    Xklij = (Vhhhh.copy()
             + ein("klic,cj->klij", Vhhhp, T1)
             + ein("klcj,ci->klij", Vhhph, T1)
             + ein("klcd,cdij->klij", Vhhpp, T2)
             + ein("klcd,ci,dj->klij", Vhhpp, T1, T1))           # :187-191
    R2 += ein("klij,abkl->abij", Xklij, T2)                      # :194
    R2 += ein("klij,ak,bl->abij", Xklij, T1, T1)                 # :197
    Xabcd = (Vpppp.copy()
             - ein("akcd,bk->abcd", Vphpp, T1)
             - ein("kbcd,ak->abcd", Vhppp, T1))                  # :200-202
    R2 += ein("abcd,cdij->abij", Xabcd, T2)                      # :205
    R2 += ein("abcd,ci,dj->abij", Xabcd, T1, T1)                 # :206

    R1 = ein("ac,ci->ai", Kac, T1) - ein("ki,ak->ai", Kki, T1)   # :215-216
    Kck = (2.0 * ein("klcd,dl->ck", Vhhpp, T1)
           - ein("kldc,dl->ck", Vhhpp, T1))                      # :219-220
    R1 += 2.0 * ein("ck,caki->ai", Kck, T2)                      # :223
    R1 -= ein("ck,caik->ai", Kck, T2)                            # :224
    R1 += ein("ck,ci,ak->ai", Kck, T1, T1)                       # :225
    R1 += 2.0 * ein("akic,ck->ai", Vphhp, T1)                    # :226
    R1 -= ein("akci,ck->ai", Vphph, T1)                          # :227
    R1 += 2.0 * ein("akcd,cdik->ai", Vphpp, T2)                  # :228
    R1 -= ein("akdc,cdik->ai", Vphpp, T2)                        # :229
    R1 += 2.0 * ein("akcd,ci,dk->ai", Vphpp, T1, T1)             # :230
    R1 -= ein("akdc,ci,dk->ai", Vphpp, T1, T1)                   # :231
    R1 -= 2.0 * ein("klic,ackl->ai", Vhhhp, T2)                  # :232
    R1 += ein("klci,ackl->ai", Vhhph, T2)                        # :233
    R1 -= 2.0 * ein("klic,ak,cl->ai", Vhhhp, T1, T1)             # :234
    R1 += ein("klci,ak,cl->ai", Vhhph, T1, T1)                   # :235
    return R1, R2
