"""FP64 CPU oracle for the cc4s CCSD amplitude iteration.

TEST INFRASTRUCTURE ONLY.  This module is the checker: it may be imported by
``tests/``, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py``, never by the product library or its Python wrapper.

It restates, in numpy (einsum -> BLAS), the reference's algorithm:

* ``slice_vertex``          SliceOperator.cxx:95-131
* ``vertex_integrals``      VertexCoulombIntegrals.cxx:55-197
* ``residuum_complex``      coupledcluster/Ccsd.cxx:381-633
* ``residuum_real``         coupledcluster/Ccsd.cxx:76-378
* ``residuum_to_amplitudes`` CoupledCluster.cxx:255-317
* ``energy``                CoupledCluster.cxx:194-253
* ``DiisMixer``             mixers/DiisMixer.cxx:27-148 (+ TensorSet.hpp:209-224)
* ``coupled_cluster``       CoupledCluster.cxx:55-192
* ``residuum_drccd``        coupledcluster/Drccd.cxx:31-109
* ``perturbative_triples``  PerturbativeTriplesReference.cxx:55-180
* ``vertex_energy`` / ``vertex_gradient``  the energy of CoupledCluster.cxx:
  194-253 as a function of Gamma (V from VertexCoulombIntegrals.cxx:68-171)
  at fixed amplitudes, its gradient and per-G split (the structure factor of
  FiniteSizeCorrection.cxx:110-124 without its 1/V(G) and singular-vector
  factors)

Arrays are indexed in the reference's index order (T2[a,b,i,j], Gamma[G,p,q])
with numpy's own memory order; only the C-ABI boundary cares about the cc4s
first-index-fastest layout.

Pinning (what makes this restatement trustworthy without the reference
binary, which needs CTF and yaml-cpp that are not available offline):
* ``oracle/ccsd_vbased.py`` restates CcsdReference.cxx:36-239, an independent
  formulation from explicit V slices (including V_pppp); tests require the two
  residua to agree to 1e-12 for real and complex inputs;
* textbook spin-orbital CCSD (Stanton, Gauss, Watts and Bartlett 1991) on the
  spin-orbital restatement of a closed-shell system converges to the same
  energy (1e-12 Eh) and amplitudes;
* the first iteration reproduces the MP2 closed form
  (SecondOrderPerturbationTheory.cxx:108-127); that second-order energy's
  singles term equals the second derivative of the exact ground-state energy
  of the one-body Hamiltonian diag(e) + lambda f (finite differences), and
  its degeneracy factors are pinned by the spin-orbital restatement of a
  closed-shell system (degeneracy 1 on the doubled orbitals gives the
  closed-shell energy);
* DIIS with maxResidua=1 is plain Jacobi; the converged fixed point is
  independent of levelShift and integralsSliceSize;
* V slices are checked against brute-force sums over G;
* the vertex gradient against central finite differences of
  ``vertex_energy`` in every real component, ``vertex_energy`` against
  ``energy`` through the V slices, and sum_G S(G) = E (degree-2
  homogeneity).
"""

from __future__ import annotations

import math
import numpy as np

try:  # LAPACK ?sysv: the same routine the reference calls (DiisMixer.cxx:129,144)
    from scipy.linalg import lapack as _lapack
except Exception:  # pragma: no cover
    _lapack = None


def ein(spec, *ops):
    return np.einsum(spec, *ops, optimize="optimal")


# --------------------------------------------------------------------------
# inputs
# --------------------------------------------------------------------------

def define_holes_and_particles(eps, fermi_energy):
    """No = number of energies below the Fermi energy
    (DefineHolesAndParticles.cxx:47-59); returns (No, eps_h, eps_p)."""
    eps = np.asarray(eps, dtype=np.float64)
    No = 0
    while No < len(eps) and eps[No] < fermi_energy:
        No += 1
    if No == 0:
        raise ValueError("Fermi energy below all eigen energies.")
    if No == len(eps):
        raise ValueError("Fermi energy above all eigen energies.")
    return No, eps[:No].copy(), eps[No:].copy()


def slice_vertex(gamma, No):
    """Cut Gamma[G,Np,Np] along both State dimensions into hh, hp, ph, pp;
    holes are [0,No), particles [Np-Nv,Np) (SliceOperator.cxx:95-131)."""
    g = np.asarray(gamma)
    return {
        "hh": g[:, :No, :No], "hp": g[:, :No, No:],
        "ph": g[:, No:, :No], "pp": g[:, No:, No:],
    }


def vertex_integrals(slices, field, names):
    """Coulomb integral slices V[p,q,s,r] named LO RO LI RI.

    complex: V = sum_G conj(Gamma_{LI LO}[G,s,p]) Gamma_{RO RI}[G,q,r]
             (VertexCoulombIntegrals.cxx:143-171)
    real:    V = sum_G Re Gamma_{LO LI}[G,p,s] Re Gamma_{RO RI}[G,q,r] + Im Im
             (VertexCoulombIntegrals.cxx:68-102)
    """
    out = {}
    for name in names:
        LO, RO, LI, RI = name
        if field == "complex":
            left = np.conj(slices[LI + LO]).transpose(0, 2, 1)   # [G,p,s]
            right = slices[RO + RI]                              # [G,q,r]
            out[name] = ein("Gps,Gqr->pqsr", left, right)
        else:
            l = slices[LO + LI]
            r = slices[RO + RI]
            out[name] = (ein("Gps,Gqr->pqsr", l.real, r.real)
                         + ein("Gps,Gqr->pqsr", l.imag, r.imag))
    return out


COMPLEX_SLICES = ("pphh", "phhp", "hhpp", "phph", "hhhh", "hhhp", "phhh")
REAL_SLICES = ("pphh", "hhhh", "hhhp")


class CcsdProblem:
    """Everything the Ccsd method reads from its step arguments."""

    def __init__(self, gamma, eps, No, field, ppl=True,
                 integrals_slice_size=None, method="Ccsd", linearized=False,
                 adjacent_pairs_exchange=False):
        if method not in ("Ccsd", "Drccd"):
            raise ValueError(f"Unknown method: '{method}'")
        self.method = method
        self.linearized = linearized
        self.adjacent_pairs_exchange = adjacent_pairs_exchange
        self.field = field
        self.No = No
        self.Nv = gamma.shape[1] - No
        self.NG = gamma.shape[0]
        self.eps_h = np.asarray(eps[:No], dtype=np.float64)
        self.eps_p = np.asarray(eps[No:], dtype=np.float64)
        self.slices = slice_vertex(gamma, No)
        names = COMPLEX_SLICES if field == "complex" else REAL_SLICES
        if method == "Drccd":
            names = ("pphh", "phhp", "hhpp")
        self.V = vertex_integrals(self.slices, field, names)
        self.ppl = ppl
        self.slice_size = integrals_slice_size or No   # Ccsd.cxx:287-289
        self.dtype = np.complex128 if field == "complex" else np.float64

    def residuum(self, amplitudes):
        if self.method == "Drccd":
            return residuum_drccd(self, amplitudes)
        if self.field == "complex":
            return residuum_complex(self, amplitudes)
        return residuum_real(self, amplitudes)


# --------------------------------------------------------------------------
# Drccd::getResiduum (both fields: the same statements, V from either rule)
# --------------------------------------------------------------------------

def residuum_drccd(P, amplitudes):
    """Drccd.cxx:31-109: direct-ring CCD.  R1 = 0 always (:61,76);
    first iteration R2 = V_abij (:62); then R2 = 2 V_akic T_cbkj (:77),
    R2_abij += R2_baji (:78), R2 += V_abij (:79) and, unless linearized,
    W_klcd = 2 V_klcd (- V_lkcd with adjacentPairsExchange),
    R2 += 2 W_klcd T_acik T_dblj (:84-99)."""
    V = P.V
    R1 = np.zeros((P.Nv, P.No), dtype=P.dtype)
    if amplitudes is None:
        return R1, V["pphh"].copy()
    T2 = amplitudes[1]
    R2 = 2.0 * ein("akic,cbkj->abij", V["phhp"], T2)
    R2 = R2 + R2.transpose(1, 0, 3, 2)
    R2 += V["pphh"]
    if not P.linearized:
        W = 2.0 * V["hhpp"]
        if P.adjacent_pairs_exchange:
            W = W - V["hhpp"].transpose(1, 0, 2, 3)
        R2 += 2.0 * ein("klcd,acik,dblj->abij", W, T2, T2)
    return R1, R2


# --------------------------------------------------------------------------
# Ccsd<Complex>::getResiduum
# --------------------------------------------------------------------------

def _ppl_slices(R2, vslab, X, s, Nv):
    """Particle-particle ladder over slice pairs m <= n
    (Ccsd.cxx:564-600 complex, 294-332 real): ``vslab(a, a+Nx, b, b+Ny)``
    returns V_xycd = sum_G left_Gxc right_Gyd for x in slice n, y in slice m;
    R_xyij = V_xycd X_cdij is added at (a,b) and, if a > b, also transposed
    at (b,a,j,i)."""
    ns = int(math.ceil(Nv / s))
    for m in range(ns):
        for n in range(m, ns):
            a, b = n * s, m * s
            x1, y1 = min(a + s, Nv), min(b + s, Nv)
            Vxycd = vslab(a, x1, b, y1)
            Rxyij = ein("xycd,cdij->xyij", Vxycd, X)
            R2[a:x1, b:y1] += Rxyij
            if a > b:
                R2[b:y1, a:x1] += Rxyij.transpose(1, 0, 3, 2)


def residuum_complex(P, amplitudes):
    """Ccsd<Complex>::getResiduum, Ccsd.cxx:381-633, statement by statement."""
    V = P.V
    Vpphh = V["pphh"]
    if amplitudes is None:                                      # :404-409
        return (np.zeros((P.Nv, P.No), dtype=np.complex128), Vpphh.copy())
    T1, T2 = amplitudes
    S = P.slices
    Gpp, Gph, Ghp, Ghh = S["pp"], S["ph"], S["hp"], S["hh"]
    cTGpp = np.conj(Gpp).transpose(0, 2, 1)                     # :432
    cTGhp = np.conj(Gph).transpose(0, 2, 1)                     # :433
    cTGph = np.conj(Ghp).transpose(0, 2, 1)                     # :434
    cTGhh = np.conj(Ghh).transpose(0, 2, 1)                     # :435
    Vphhp, Vhhpp, Vphph = V["phhp"], V["hhpp"], V["phph"]
    Vhhhh, Vhhhp, Vphhh = V["hhhh"], V["hhhp"], V["phhh"]

    T1T1 = ein("ai,bj->abij", T1, T1)
    X = T2 + T1T1                                               # :460-461
    Y = T2 + 2.0 * T1T1                                         # :462-463
    Kac = (-2.0 * ein("klcd,adkl->ac", Vhhpp, X)
           + ein("kldc,adkl->ac", Vhhpp, X))                    # :465-466
    Lac = (Kac + 2.0 * ein("Gac,Gkd,dk->ac", cTGpp, Ghp, T1)
           - ein("Gad,Gkc,dk->ac", cTGpp, Ghp, T1))             # :468-470
    Kki = (2.0 * ein("klcd,cdil->ki", Vhhpp, X)
           - ein("kldc,cdil->ki", Vhhpp, X))                    # :472-473
    Lki = (Kki + 2.0 * ein("klic,cl->ki", Vhhhp, T1)
           - ein("lkic,cl->ki", Vhhhp, T1))                     # :476-478
    R2 = ein("ac,cbij->abij", Lac, T2)                          # :481
    R2 -= ein("ki,abkj->abij", Lki, T2)                         # :484
    cTDGph = cTGph - ein("Gki,ak->Gai", cTGhh, T1)              # :487-488
    R2 += ein("Gai,Gbc,cj->abij", cTDGph, Gpp, T1)              # :489
    R2 -= ein("akij,bk->abij", Vphhh, T1)                       # :490
    R2 -= ein("akic,cj,bk->abij", Vphhp, T1, T1)                # :491
    cTDGph = cTDGph + ein("Gad,di->Gai", cTGpp, T1)             # :494
    cTDGph = cTDGph - 0.5 * ein("Gld,dail->Gai", cTGhp, Y)      # :495
    Xakic = ein("Gai,Gkc->akic", cTDGph, Ghp)                   # :496
    # Y[d,c,l,k] = Vhhpp[l,k,d,c] - 1/2 Vhhpp[l,k,c,d]           # :497-498
    Yv = (ein("lkdc->dclk", Vhhpp) - 0.5 * ein("lkcd->dclk", Vhhpp))
    Xakic = Xakic + ein("dclk,adil->akic", Yv, T2)              # :499
    Yt = 2.0 * T2 - T2.transpose(1, 0, 2, 3)                    # :500-501
    R2 += ein("akic,cbkj->abij", Xakic, Yt)                     # :502
    cTDGpp = cTGpp - ein("Glc,al->Gac", cTGhp, T1)              # :509-510
    dGhh = Ghh + ein("Gkd,di->Gki", Ghp, T1)                    # :511-512
    Xakci = ein("Gac,Gki->akci", cTDGpp, dGhh)                  # :514
    Xakci = Xakci - 0.5 * ein("lkcd,dail->akci", Vhhpp, T2)     # :515
    R2 -= ein("akci,cbkj->abij", Xakci, T2)                     # :516
    R2 -= ein("bkci,ackj->abij", Xakci, T2)                     # :517
    R2 = R2 + R2.transpose(1, 0, 3, 2)                          # :520
    R2 += Vpphh                                                 # :529
    Xklij = (Vhhhh + ein("klic,cj->klij", Vhhhp, T1)
             + ein("lkjc,ci->klij", Vhhhp, T1))                 # :534-536
    R2 += ein("klij,abkl->abij", Xklij, X)                      # :538
    Xklij = ein("klcd,cdij->klij", Vhhpp, X)                    # :541
    R2 += ein("klij,abkl->abij", Xklij, T2 if P.ppl else X)     # :542-546
    if P.ppl:                                                   # :549-600
        dGpp = Gpp - ein("Gkb,ak->Gab", Ghp, T1)                # :560-561

        def vslab(a, x1, b, y1):                                # :588-589
            return ein("Gxc,Gyd->xycd", cTDGpp[:, a:x1, :], dGpp[:, b:y1, :])
        _ppl_slices(R2, vslab, X, P.slice_size, P.Nv)

    R1 = ein("ac,ci->ai", Kac, T1) - ein("ki,ak->ai", Kki, T1)  # :609-610
    Kck = (2.0 * ein("klcd,dl->ck", Vhhpp, T1)
           - ein("kldc,dl->ck", Vhhpp, T1))                     # :613-614
    R1 += 2.0 * ein("ck,caki->ai", Kck, T2)                     # :617
    R1 -= ein("ck,caik->ai", Kck, T2)                           # :618
    R1 += ein("ck,ci,ak->ai", Kck, T1, T1)                      # :619
    R1 += 2.0 * ein("akic,ck->ai", Vphhp, T1)                   # :620
    R1 -= ein("akci,ck->ai", Vphph, T1)                         # :621
    R1 += 2.0 * ein("Gac,Gkd,cdik->ai", cTGpp, Ghp, X)          # :622-623
    R1 -= ein("Gad,Gkc,cdik->ai", cTGpp, Ghp, X)                # :624-625
    R1 += (-2.0 * ein("klic,ackl->ai", Vhhhp, X)
           + ein("lkic,ackl->ai", Vhhhp, X))                    # :627-628
    return R1, R2


# --------------------------------------------------------------------------
# Ccsd<Real>::getResiduum
# --------------------------------------------------------------------------

def residuum_real(P, amplitudes):
    """Ccsd<Real>::getResiduum, Ccsd.cxx:76-378, statement by statement."""
    V = P.V
    Vpphh = V["pphh"]
    if amplitudes is None:                                      # :97-102
        return (np.zeros((P.Nv, P.No)), Vpphh.copy())
    T1, T2 = amplitudes
    S = P.slices
    rGpp, iGpp = S["pp"].real, S["pp"].imag                     # :129-136
    rGph, iGph = S["ph"].real, S["ph"].imag
    rGhh, iGhh = S["hh"].real, S["hh"].imag
    Vhhhh, Vhhhp = V["hhhh"], V["hhhp"]

    X = T2 + ein("ai,bj->abij", T1, T1)                         # :169-170
    Kki = (2.0 * ein("cdkl,cdil->ki", Vpphh, X)
           - ein("dckl,cdil->ki", Vpphh, X))                    # :172-173
    Lki = Kki.copy()                                            # :175
    Kac = (-2.0 * ein("cdkl,adkl->ac", Vpphh, X)
           + ein("dckl,adkl->ac", Vpphh, X))                    # :177-178
    Lac = (Kac
           + 2.0 * ein("Gca,Gdk,dk->ac", rGpp, rGph, T1)
           + 2.0 * ein("Gca,Gdk,dk->ac", iGpp, iGph, T1)
           - ein("Gck,Gda,dk->ac", rGph, rGpp, T1)
           - ein("Gck,Gda,dk->ac", iGph, iGpp, T1))             # :180-189
    Lki += (2.0 * ein("klic,cl->ki", Vhhhp, T1)
            - ein("lkic,cl->ki", Vhhhp, T1))                    # :192-193
    R2 = ein("ac,cbij->abij", Lac, T2)                          # :195
    R2 -= ein("ki,abkj->abij", Lki, T2)                         # :197
    rDGph = rGph - ein("Gki,ak->Gai", rGhh, T1)                 # :201-202
    iDGph = iGph - ein("Gki,ak->Gai", iGhh, T1)                 # :203-204
    R2 += ein("Gai,Gbc,cj->abij", rDGph, rGpp, T1)              # :205-206
    R2 += ein("Gai,Gbc,cj->abij", iDGph, iGpp, T1)              # :207-208
    R2 -= ein("jika,bk->abij", Vhhhp, T1)                       # :209
    R2 -= ein("bk,acik,cj->abij", T1, Vpphh, T1)                # :210
    X = X + ein("ai,bj->abij", T1, T1)                          # :215
    Xakic = -0.5 * ein("dclk,dail->akic", Vpphh, X)             # :216
    Xakic += ein("dclk,adil->akic", Vpphh, T2)                  # :217
    Xakic -= 0.5 * ein("cdlk,adil->akic", Vpphh, T2)            # :218
    rDGph = rDGph + ein("Gad,di->Gai", rGpp, T1)                # :220
    iDGph = iDGph + ein("Gad,di->Gai", iGpp, T1)                # :221
    Xakic += ein("Gai,Gck->akic", rDGph, rGph)                  # :223
    Xakic += ein("Gai,Gck->akic", iDGph, iGph)                  # :224
    X = 2.0 * T2 - T2.transpose(1, 0, 2, 3)                     # :227-228
    R2 += ein("akic,cbkj->abij", Xakic, X)                      # :229
    Xakci = -0.5 * ein("cdlk,dail->akci", Vpphh, T2)            # :239
    rDGpp = rGpp - ein("Gcl,al->Gac", rGph, T1)                 # :240,242
    iDGpp = iGpp - ein("Gcl,al->Gac", iGph, T1)                 # :241,243
    rDGhh = rGhh + ein("Gdk,di->Gki", rGph, T1)                 # :244,246
    iDGhh = iGhh + ein("Gdk,di->Gki", iGph, T1)                 # :245,247
    Xakci += ein("Gac,Gki->akci", rDGpp, rDGhh)                 # :248
    Xakci += ein("Gac,Gki->akci", iDGpp, iDGhh)                 # :249
    R2 -= ein("akci,cbkj->abij", Xakci, T2)                     # :252
    R2 -= ein("bkci,ackj->abij", Xakci, T2)                     # :253
    R2 = R2 + R2.transpose(1, 0, 3, 2)                          # :256
    R2 += Vpphh                                                 # :263
    X = T2 + ein("ai,bj->abij", T1, T1)                         # :267-268
    Xklij = (Vhhhh + ein("klic,cj->klij", Vhhhp, T1)
             + ein("lkjc,ci->klij", Vhhhp, T1))                 # :269-271
    R2 += ein("klij,abkl->abij", Xklij, X)                      # :273
    Xklij = ein("cdkl,cdij->klij", Vpphh, X)                    # :275
    R2 += ein("klij,abkl->abij", Xklij, T2 if P.ppl else X)     # :276-280
    if P.ppl:                                                   # :283-333
        def vslab(a, x1, b, y1):                                # :318-321
            return (ein("Gxc,Gyd->xycd", rDGpp[:, a:x1, :], rDGpp[:, b:y1, :])
                    + ein("Gxc,Gyd->xycd", iDGpp[:, a:x1, :],
                          iDGpp[:, b:y1, :]))
        _ppl_slices(R2, vslab, X, P.slice_size, P.Nv)

    R1 = ein("ac,ci->ai", Kac, T1) - ein("ki,ak->ai", Kki, T1)  # :341-342
    Kck = (2.0 * ein("cdkl,dl->ck", Vpphh, T1)
           - ein("dckl,dl->ck", Vpphh, T1))                     # :345-346
    R1 += 2.0 * ein("ck,caki->ai", Kck, T2)                     # :349
    R1 -= ein("ck,caik->ai", Kck, T2)                           # :350
    R1 += ein("ak,ck,ci->ai", T1, Kck, T1)                      # :351
    R1 += 2.0 * ein("acik,ck->ai", Vpphh, T1)                   # :352
    R1 -= ein("Gca,Gik,ck->ai", rGpp, rGhh, T1)                 # :358-359
    R1 -= ein("Gca,Gik,ck->ai", iGpp, iGhh, T1)                 # :360-361
    R1 += 2.0 * ein("Gca,Gdk,cdik->ai", rGpp, rGph, X)          # :363-364
    R1 += 2.0 * ein("Gca,Gdk,cdik->ai", iGpp, iGph, X)          # :365-366
    R1 -= ein("Gda,Gck,cdik->ai", rGpp, rGph, X)                # :367-368
    R1 -= ein("Gda,Gck,cdik->ai", iGpp, iGph, X)                # :369-370
    R1 += (-2.0 * ein("klic,ackl->ai", Vhhhp, X)
           + ein("lkic,ackl->ai", Vhhhp, X))                    # :372-373
    return R1, R2


# --------------------------------------------------------------------------
# CoupledCluster: denominators, energy, DIIS, loop
# --------------------------------------------------------------------------

def energy_differences(eps_h, eps_p, rank):
    """Delta_{a..i..} = sum_p eps_p - eps_h, accumulated as
    D += eps_p[a]; D -= eps_h[i] per excitation (CoupledCluster.cxx:303-316)."""
    if rank == 1:
        return (0.0 + eps_p[:, None]) - eps_h[None, :]
    D = np.zeros((len(eps_p), len(eps_p), len(eps_h), len(eps_h)))
    D += eps_p[:, None, None, None]
    D -= eps_h[None, None, :, None]
    D += eps_p[None, :, None, None]
    D -= eps_h[None, None, None, :]
    return D


def residuum_to_amplitudes(P, residuum, amplitudes, level_shift=0.0):
    """T' = R * (-1/(Delta + s)), with R -= s*T first when amplitudes are
    given and s != 0 (CoupledCluster.cxx:255-283)."""
    R1, R2 = residuum
    if amplitudes is not None and level_shift != 0.0:           # :264-267
        R1 = R1 - level_shift * amplitudes[0]
        R2 = R2 - level_shift * amplitudes[1]
    D1 = -1.0 / (energy_differences(P.eps_h, P.eps_p, 1) + level_shift)
    D2 = -1.0 / (energy_differences(P.eps_h, P.eps_p, 2) + level_shift)
    return R1 * D1, R2 * D2                                     # :275-281


def energy(P, amplitudes):
    """(E, D, X): D = 2 sum tau_abij conj V_abij, X = -sum tau_abij conj V_baij,
    tau = T2 + T1 T1 (CoupledCluster.cxx:204-235)."""
    T1, T2 = amplitudes
    Vc = np.conj(P.V["pphh"])                                   # :206-208
    deg = 2.0                                                   # :212
    D = (0.5 * deg * deg * ein("abij,abij->", T2, Vc)
         + 0.5 * deg * deg * ein("ai,bj,abij->", T1, T1, Vc))   # :224-227
    Xc = Vc.transpose(1, 0, 2, 3)
    X = (-0.5 * deg * ein("abij,abij->", T2, Xc)
         - 0.5 * deg * ein("ai,bj,abij->", T1, T1, Xc))         # :228-231
    return D + X, D, X


def tensor_set_dot(x, y):
    """<x|y> = sum over components of conj(x) * y (TensorSet.hpp:209-224)."""
    return sum(np.vdot(a.ravel(), b.ravel()) for a, b in zip(x, y))


class DiisMixer:
    """DIIS over (amplitudes, amplitude change) pairs (DiisMixer.cxx:27-148)."""

    def __init__(self, max_residua=4, complex_field=True):
        self.N = N = int(max_residua)                           # :33
        self.amplitudes = [None] * N
        self.residua = [None] * N
        self.next_index = 0
        self.count = 0
        M = N + 1
        self.B = np.zeros(M * M)                                # :42-44
        self.B[1:M] = -1.0
        self.B[M:M * M:M] = -1.0
        self.complex_field = complex_field
        self.next = None
        self.residuum_norm = None
        self.last_coefficients = None

    def _inverse(self, matrix, dim):
        """Solve matrix * c = (-1,0,...) with ?sysv('U') (:119-148)."""
        rhs = np.zeros(dim)
        rhs[0] = -1.0
        A = matrix.reshape(dim, dim)          # row m, column n = B[n + dim*m]
        if _lapack is not None:
            if self.complex_field:
                A = A.astype(np.complex128)
                rhs = rhs.astype(np.complex128)
                out = _lapack.zsysv(A.T, rhs, lower=0)
            else:
                out = _lapack.dsysv(A.T, rhs, lower=0)
            info = out[-1]
            if info != 0:
                raise RuntimeError("Diagonalization failed")
            return out[2]
        return np.linalg.solve(A.T, rhs)

    def append(self, A, R):
        N = self.N
        self.amplitudes[self.next_index] = A                    # :65-66
        self.residua[self.next_index] = R
        for i in range(N):                                      # :69-77
            if self.residua[i] is not None:
                overlap = 2.0 * tensor_set_dot(self.residua[i], R).real
                self.B[(i + 1) * (N + 1) + self.next_index + 1] = overlap
                self.B[(self.next_index + 1) * (N + 1) + i + 1] = overlap
        if self.count < N:                                      # :80
            self.count += 1
        dim = self.count + 1
        matrix = np.empty(dim * dim)
        for m in range(dim):
            for n in range(dim):
                matrix[n + dim * m] = self.B[n + (N + 1) * m]   # :84-86
        column = self._inverse(matrix, dim)                     # :89
        nxt = [np.zeros_like(c) for c in A]                     # :92-95
        nres = [np.zeros_like(c) for c in R]
        coeffs = []
        for j in range(self.count):                             # :97-103
            i = (self.next_index + N - j) % N
            c = column[i + 1]
            coeffs.append(c)
            for k in range(len(nxt)):
                nxt[k] = nxt[k] + c * self.amplitudes[i][k]
                nres[k] = nres[k] + c * self.residua[i][k]
        self.last_coefficients = coeffs
        self.next_index = (self.next_index + 1) % N             # :105
        self.next = tuple(nxt)
        self.residuum_norm = math.sqrt(tensor_set_dot(nres, nres).real)  # :106

    def get(self):
        return self.next

    def get_residuum_norm(self):
        return self.residuum_norm


def coupled_cluster(P, max_iterations=16, energy_convergence=1e-7,
                    amplitudes_convergence=1e-6, level_shift=0.0,
                    max_residua=4, initial_amplitudes=None, callback=None):
    """CoupledCluster::run (CoupledCluster.cxx:55-192) with the Ccsd method
    and the DiisMixer.  Defaults are the reference's (CoupledCluster.hpp:45-49,
    DiisMixer.cxx:33)."""
    amplitudes = initial_amplitudes
    mixer = DiisMixer(max_residua, complex_field=(P.field == "complex"))
    e = 0.0
    previous_e = 0.0
    result = {"energy": {}, "trace": []}
    i = 0
    while i < max_iterations:                                   # :129
        residuum = P.residuum(amplitudes)                       # :136
        new = residuum_to_amplitudes(P, residuum, amplitudes, level_shift)
        if amplitudes is not None:                              # :138-144
            change = tuple(n - a for n, a in zip(new, amplitudes))
            second_order = False
        else:
            change = tuple(np.copy(n) for n in new)
            second_order = True
        mixer.append(new, change)                               # :145
        amplitudes = mixer.get()                                # :147
        dR = mixer.get_residuum_norm()                          # :148
        e, _, _ = energy(P, amplitudes)                         # :149
        result["trace"].append((float(np.real(e)),
                                float(np.real(e - previous_e)), dR))
        if callback:
            callback(i, e, dR)
        if second_order:                                        # :164-166
            result["energy"]["secondOrder"] = float(np.real(e))
        if abs(e - previous_e) < energy_convergence and \
                dR < amplitudes_convergence:                    # :167-173
            break
        previous_e = e
        i += 1
    if amplitudes is None:
        raise ValueError("no amplitudes: maxIterations 0 needs initial ones")
    e, D, X = energy(P, amplitudes)                             # :184
    result["energy"].update({"correlation": float(np.real(e)),
                             "direct": float(np.real(D)),
                             "exchange": float(np.real(X)),
                             "unit": 1.0})
    result["convergenceReached"] = i < max_iterations           # :185
    result["iterations"] = min(i + 1, max_iterations)
    result["amplitudes"] = amplitudes
    return result


def second_order(P, fph=None, degeneracy=2.0):
    """SecondOrderPerturbationTheory::run (SecondOrderPerturbationTheory.cxx:
    54-149): first-order amplitudes T_ai = conj(f_ai) / (e_a - e_i) and
    T_abij = conj(V_abij) / ((e_a - e_i) + (e_b - e_j)) (:98-120); singles
    S = -g sum f_ai T_ai, direct D = -g^2/2 sum V_abij T_abij, exchange
    X = +g/2 sum V_abji T_abij (:121-127) with the spin degeneracy g (2 for
    closed shells, 1 when fewer states carry the first spin value than there
    are states, :67-85).  fph [a, i]: the particle-hole block of the Fock
    operator (None: canonical orbitals, f_ph = 0, :155-185).  Returns the
    real parts {singles, direct, exchange, correlation}."""
    g = float(degeneracy)
    V = P.V["pphh"]
    dph = P.eps_p[:, None] - P.eps_h[None, :]
    Dpphh = dph[:, None, :, None] + dph[None, :, None, :]
    T = np.conj(V) * (1.0 / Dpphh)
    S = 0.0
    if fph is not None:
        S = -g * np.sum(fph * (np.conj(fph) * (1.0 / dph)))
    D = -0.5 * g * g * np.sum(V * T)
    X = 0.5 * g * np.sum(V.transpose(0, 1, 3, 2) * T)
    out = {"singles": float(np.real(S)), "direct": float(np.real(D)),
           "exchange": float(np.real(X))}
    out["correlation"] = float(np.real(S + D + X))
    return out


def mp2_energy(P):
    """Second-order correlation energy of canonical closed-shell orbitals
    (f_ph = 0, degeneracy 2), SecondOrderPerturbationTheory.cxx:108-127."""
    return second_order(P)["correlation"]


# ---------------------------------------------------------------------------
# Energy as a function of the vertex (adjoint checks)
# ---------------------------------------------------------------------------

def energy_weights(amplitudes, field):
    """Z with E = sum_abij Z_abij conj V_abij: Z = 2 tau_abij - tau_baij,
    tau = T2 + T1 T1 (CoupledCluster.cxx:204-232).  Real field: the
    symmetrised Z_abij + Z_baji (V_abij is formed from one vertex block
    twice there, VertexCoulombIntegrals.cxx:87-102)."""
    T1, T2 = amplitudes
    tau = T2 + ein("ai,bj->abij", T1, T1)
    Z = 2.0 * tau - tau.transpose(1, 0, 2, 3)
    if field == "real":
        Z = Z + Z.transpose(1, 0, 3, 2)
    return Z


def vertex_energy(gamma, amplitudes, No, field):
    """Re E (CoupledCluster.cxx:194-253) with V_abij formed from `gamma`
    (VertexCoulombIntegrals.cxx:68-102 real, 143-171 complex)."""
    V = vertex_integrals(slice_vertex(gamma, No), field, ("pphh",))["pphh"]
    T1, T2 = amplitudes
    tau = T2 + ein("ai,bj->abij", T1, T1)
    Vc = np.conj(V)
    E = 2.0 * np.sum(tau * Vc) - np.sum(tau * Vc.transpose(1, 0, 2, 3))
    return float(np.real(E))


def vertex_gradient(gamma, amplitudes, No, field):
    """(g, S): g[G,p,q] = dE/dRe Gamma^G_pq + i dE/dIm Gamma^G_pq of
    ``vertex_energy`` and S[G] = 1/2 sum_pq Re(conj(Gamma^G_pq) g^G_pq).

    complex: E = sum_G sum_abij Z_abij Gamma^G_ia conj(Gamma^G_bj), so
      g^G_ia = sum_bj conj(Z_abij) Gamma^G_bj, g^G_bj = sum_ai Z_abij Gamma^G_ia;
    real: E = 1/2 sum_G sum_abij Zs_abij (Re Re + Im Im)(Gamma^G_ai, Gamma^G_bj),
      g^G_ai = sum_bj Zs_abij Gamma^G_bj.
    The hh and pp blocks do not enter E."""
    g = np.zeros(gamma.shape, dtype=np.complex128)
    Z = energy_weights(amplitudes, field)
    hp = gamma[:, :No, No:]
    ph = gamma[:, No:, :No]
    if field == "complex":
        g[:, :No, No:] = ein("abij,Gbj->Gia", np.conj(Z), ph)
        g[:, No:, :No] = ein("abij,Gia->Gbj", Z, hp)
    else:
        g[:, No:, :No] = ein("abij,Gbj->Gai", Z, ph)
    S = 0.5 * np.real(np.sum(np.conj(gamma) * g, axis=(1, 2)))
    return g, S


# ---------------------------------------------------------------------------
# PerturbativeTriplesReference (the (T) energy of converged amplitudes)
# ---------------------------------------------------------------------------

def mp2_pair_energies(P):
    """MP2 pair energies of the finite basis as PerturbativeTriplesStar.hpp:
    74-88 forms them: M_abij = conj(V_abij) / (e_i + e_j - e_a - e_b),
    e_ij = sum_ab 2 M_abij V_abij - M_abij V_abji (complex, [i, j])."""
    V = P.V["pphh"]
    dph = P.eps_p[:, None] - P.eps_h[None, :]
    M = np.conj(V) / -(dph[:, None, :, None] + dph[None, :, None, :])
    return ein("abij,abij->ij", M, 2.0 * V - V.transpose(0, 1, 3, 2))


def triples_star(P, e_triples, pair_energies_cbs):
    """The (T*) correction of PerturbativeTriplesStar.hpp:34-103: the (T)
    energy scaled by sum_ij of the given (basis-set limit) MP2 pair energies
    over the finite-basis MP2 energy sum_ij e_ij."""
    ratio = float(np.sum(pair_energies_cbs)) / float(np.real(np.sum(mp2_pair_energies(P))))
    return ratio * e_triples


def perturbative_triples(gamma, eps, No, field, amplitudes, invariant=False):
    """(T) correlation energy as PerturbativeTriplesReference.cxx:55-180
    forms it (cT off):
      T'_abcijk = V_bcdk T_adij - conj(V_jklc) T_abil             (:106-107)
      W = T' + T'_bacjik + T'_acbikj + T'_cbakji + T'_cabkij + T'_bcajki
      S = W + T1_ai V_bcjk + T1_bj V_acik + T1_ck V_abij          (:118-121)
      Z = 4/3 S_abcijk - 2 S_acbijk + 2/3 S_bcaijk                 (:123-125)
      D = e_i + e_j + e_k - e_a - e_b - e_c                        (:128-134)
      E = sum conj(W)/D * conj(Z)/D * D = sum conj(W Z) / D       (:136-140,163)
    with V_ppph, V_hhhp, V_pphh from VertexCoulombIntegrals.  Returns Re E.

    W and Z both pick up conj(u_a u_b u_c) u_i u_j u_k under an orbital
    phase change (Gamma^G_pq -> conj(u_p) Gamma^G_pq u_q), so the reference's
    conj(W Z) changes with the phases whenever V is not real.  invariant=True
    returns Re sum conj(W) Z / D instead, which does not change and equals
    the reference's value for real V.  Memory O(o^3 v^3): small cases only."""
    W, Z, D = triples_parts(gamma, eps, No, field, amplitudes)
    if invariant:
        return float(np.real(np.sum(np.conj(W) * Z / D)))
    return float(np.real(np.sum(np.conj(W * Z) / D)))


def triples_parts(gamma, eps, No, field, amplitudes):
    """(W, Z, D) of ``perturbative_triples`` as [a,b,c,i,j,k] arrays."""
    slices = slice_vertex(gamma, No)
    V = vertex_integrals(slices, field, ("pphh", "ppph", "hhhp"))
    T1, T2 = amplitudes
    Tp = (ein("bcdk,adij->abcijk", V["ppph"], T2)
          - ein("jklc,abil->abcijk", np.conj(V["hhhp"]), T2))
    W = (Tp + Tp.transpose(1, 0, 2, 4, 3, 5) + Tp.transpose(0, 2, 1, 3, 5, 4)
         + Tp.transpose(2, 1, 0, 5, 4, 3) + Tp.transpose(1, 2, 0, 4, 5, 3)
         + Tp.transpose(2, 0, 1, 5, 3, 4))
    R = (ein("ai,bcjk->abcijk", T1, V["pphh"])
         + ein("bj,acik->abcijk", T1, V["pphh"])
         + ein("ck,abij->abcijk", T1, V["pphh"]))
    S = W + R
    Z = (4.0 / 3.0) * S - 2.0 * S.transpose(0, 2, 1, 3, 4, 5) \
        + (2.0 / 3.0) * S.transpose(2, 0, 1, 3, 4, 5)
    eh = np.asarray(eps[:No], dtype=np.float64)
    ep = np.asarray(eps[No:], dtype=np.float64)
    D = (eh[None, None, None, :, None, None] + eh[None, None, None, None, :, None]
         + eh[None, None, None, None, None, :] - ep[:, None, None, None, None, None]
         - ep[None, :, None, None, None, None] - ep[None, None, :, None, None, None])
    return W, Z, D
