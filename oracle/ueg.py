"""Uniform electron gas inputs, restated from UegVertexGenerator.cxx.

TEST INFRASTRUCTURE ONLY (see oracle/ccsd_oracle.py): imported by tests/ only.

The reference builds a plane-wave Coulomb vertex and Hartree-Fock orbital
energies for a closed-shell uniform electron gas (UegVertexGenerator.cxx:
80-297).  Its test `test/tests/ueg/rs1.0-7occ-26virt` runs Drccd, Ccsd and
(T) on the generator's output for rs = 1, No = 7, Nv = 26 and records the
energies in `correct.out.yaml`; the input files themselves are remote, but
restating the generator reproduces them, so those recorded energies become a
known-answer test of the whole CCSD path (the test's own tolerance is 1e-7,
test/tests/ueg/rs1.0-7occ-26virt/check.py).

Arrays: Gamma[G, p, q] (NG, Np, Np) complex, Fortran order as cc4s stores
tensors; eps[p].
"""

from __future__ import annotations

import itertools
import math

import numpy as np


def madelung(v: float) -> float:
    """UegVertexGenerator::evalMadelung (UegVertexGenerator.cxx:27-45): Ewald
    sums over l in [-6, 6]^3 for the cell volume v, as written there."""
    kappa = v ** (-1.0 / 3.0)
    term2 = math.pi / (kappa * kappa * v)
    term4 = 2.0 * kappa / math.sqrt(math.pi)
    box = 1.0 / kappa
    recip = 0.0
    real = 0.0
    for l1 in range(-6, 7):
        for l2 in range(-6, 7):
            for l3 in range(-6, 7):
                n2 = l1 * l1 + l2 * l2 + l3 * l3
                if n2 == 0:
                    continue
                modr = box * math.sqrt(n2)
                k2 = kappa * kappa * n2
                recip -= 1.0 / (math.pi * k2) * math.exp(-math.pi * math.pi * k2 / kappa / kappa) / v
                real -= math.erfc(kappa * modr) / modr
    return real + term2 + term4 + recip


def generate(No: int, Nv: int, rs: float, hartree_fock: bool = True):
    """(gamma, eps, fermi_energy) of UegVertexGenerator::run with NF = 0 (the
    full momentum-transfer grid), halfGrid = 0 and the default madelung.

    States: integer grid points g in [-maxG, maxG]^3, maxG = floor((5 Np)^(1/3)),
    sorted by |g|^2 (:107-116); the order inside a shell is irrelevant to every
    energy (it permutes degenerate states consistently).  Volume
    v = 4/3 pi rs^3 * 2 No, b = 2 pi / v^(1/3) (:129-132).  Orbital energies
    |k|^2/2 minus the exchange sum over occupied states with
    V(k, k') = 4 pi / (v |k - k'|^2), madelung for k = k' (:46-51, :141-148).
    Vertex element (G, p, q) = sqrt(4 pi / (v b^2 |g_p - g_q|^2)) at the
    momentum index G of g_p - g_q, sqrt(madelung) for p = q (:213-296; the
    reference writes index ii + NF*(q + Np*p) with d = g_q - g_p, i.e.
    element (G, q, p) of d = g_q - g_p)."""
    Np = No + Nv
    maxG = int((5.0 * Np) ** (1.0 / 3.0))
    grid = [g for g in itertools.product(range(-maxG, maxG + 1), repeat=3)]
    grid.sort(key=lambda g: g[0] * g[0] + g[1] * g[1] + g[2] * g[2])
    sl = [g[0] * g[0] + g[1] * g[1] + g[2] * g[2] for g in grid]
    if sl[No] == sl[No - 1]:
        raise ValueError("occupied orbitals do not form a closed shell")
    grid = np.array(grid[:Np], dtype=np.int64)

    v = rs ** 3 / 3.0 * 4.0 * math.pi * No * 2
    a = v ** (1.0 / 3.0)
    b = 2.0 * math.pi / a
    mad = madelung(v)

    k = b * grid.astype(np.float64)
    eps = np.empty(Np)
    for p in range(Np):
        e = 0.5 * float(k[p] @ k[p])
        exch = 0.0
        for o in range(No):
            q = k[p] - k[o]
            q2 = float(q @ q)
            exch += mad if q2 < 1e-8 else 4.0 * math.pi / v / q2
        eps[p] = e - exch if hartree_fock else e
    fermi = 0.5 * (eps[No] + eps[No - 1])

    # momentum-transfer grid: all integer vectors with |t|^2 <= max |g_p - g_q|^2
    diffs = grid[:, None, :] - grid[None, :, :]
    d2 = np.sum(diffs * diffs, axis=2)
    maxR = int(d2.max())
    maxT = int(np.max(np.abs(diffs[d2 == maxR])))
    index = {}
    for t in itertools.product(range(-maxT, maxT + 1), repeat=3):
        if t[0] * t[0] + t[1] * t[1] + t[2] * t[2] <= maxR:
            index[t] = len(index)
    NG = len(index)
    fac = 4.0 * math.pi / v
    gamma = np.zeros((NG, Np, Np), dtype=np.complex128, order="F")
    for p in range(Np):
        for q in range(Np):
            d = tuple(int(x) for x in grid[p] - grid[q])
            n2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
            res = fac / (b * b * n2) if n2 else mad
            gamma[index[d], p, q] = math.sqrt(res)
    return gamma, eps, fermi


def momentum_grid(No: int, Nv: int, rs: float):
    """(grid [NG, 3], B [3, 3], V [NG]) of the vertex that `generate` builds:
    the momentum transfer b t of each G index, the reciprocal lattice vectors
    b e_x, b e_y, b e_z, and the Coulomb potential 4 pi / (v |b t|^2) (the
    madelung term at t = 0), i.e. Gamma^G_pq^2 for the pairs with momentum
    transfer t (UegVertexGenerator.cxx:213-296)."""
    Np = No + Nv
    maxG = int((5.0 * Np) ** (1.0 / 3.0))
    grid = [g for g in itertools.product(range(-maxG, maxG + 1), repeat=3)]
    grid.sort(key=lambda g: g[0] * g[0] + g[1] * g[1] + g[2] * g[2])
    grid = np.array(grid[:Np], dtype=np.int64)
    v = rs ** 3 / 3.0 * 4.0 * math.pi * No * 2
    b = 2.0 * math.pi / v ** (1.0 / 3.0)
    diffs = grid[:, None, :] - grid[None, :, :]
    d2 = np.sum(diffs * diffs, axis=2)
    maxR = int(d2.max())
    maxT = int(np.max(np.abs(diffs[d2 == maxR])))
    ts = [t for t in itertools.product(range(-maxT, maxT + 1), repeat=3)
          if t[0] * t[0] + t[1] * t[1] + t[2] * t[2] <= maxR]
    fac = 4.0 * math.pi / v
    pot = np.array([fac / (b * b * (t[0] * t[0] + t[1] * t[1] + t[2] * t[2]))
                    if any(t) else madelung(v) for t in ts])
    return b * np.array(ts, dtype=np.float64), b * np.eye(3), pot


# recorded by the reference for rs = 1, No = 7, Nv = 26
# (test/tests/ueg/rs1.0-7occ-26virt/correct.out.yaml:120-170)
UEG_RS1_NO7_NV26 = {
    "Drccd": {"correlation": -0.2758298986953282, "direct": -0.4128084206966558,
              "exchange": 0.1369785220013276, "secondOrder": -0.36143028565969504},
    "Ccsd": {"correlation": -0.39269658954585018, "direct": -0.5953888037963122,
             "exchange": 0.20269221425046202, "secondOrder": -0.36143028565969504},
    # PerturbativeTriplesReference on the Ccsd amplitudes (:166-170)
    "Triples": {"correlation": -0.0063019625641725016},
    "settings": {"maxIterations": 20, "energyConvergence": 1e-8,
                 "amplitudesConvergence": 1e-8, "maxResidua": 4,
                 "integralsSliceSize": 100},
}
