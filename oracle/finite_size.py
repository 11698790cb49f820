"""FP64 CPU restatement of FiniteSizeCorrection (algorithms/FiniteSizeCorrection.cxx).

TEST INFRASTRUCTURE ONLY (see oracle/ccsd_oracle.py): imported by tests/ only.

Two parts:

* the transition structure factor at given amplitudes
  (FiniteSizeCorrection.cxx:49-133): the vertex is brought to the momentum
  grid with the singular vectors, Gamma^G = sum_F Gamma^F U[G,F] (:71-80),
  divided by sqrt(V(G)) into codensities C (:87-102), and
  S(G) = Re[ 2 cTC_Gia C_Gjb tau_abij - cTC_Gja C_Gib tau_abij ] with
  cTC_Gia = conj(C_Gai) (ph block) and tau = T2 + T1 T1 (:111-125);
* the interpolation of S on a finer grid and the integration of
  S(G)/|G|^2 (:136-311): S is placed on the bounding box of the grid in
  direct (reciprocal-lattice) coordinates, interpolated with clamped
  Catmull-Rom tricubic splines (the vendored Geometric Tools
  IntpTricubic3, gte/TricubicInterpolation.hpp:231-353, catmullRom true),
  and summed over the (2N)^3 shifts of every grid point.

Parity pins (tests/test_finite_size.py): the spline reproduces the grid
values and quadratic functions exactly; with identity singular vectors,
sum_G V(G) S(G) equals the CCSD correlation energy of the same amplitudes
(the energy decomposed over G); for a grid filling its bounding box and a
constant S the corrected energy is the plain lattice sum over the fine grid,
the vertex unit cancels and the result is linear in S; the host restatement
in the library is compared to this one on synthetic grids.
"""

from __future__ import annotations

import math

import numpy as np


def transition_structure_factor(gamma, No, T1, T2, V, U=None):
    """S[G] (real) of FiniteSizeCorrection.cxx:49-133.  gamma [NF, Np, Np]
    complex (the sliced vertex, holes first), T1 [a,i], T2 [a,b,i,j],
    V [NG] the Coulomb potential, U [NG, NF] the singular vectors (None: the
    vertex is already on the grid)."""
    ph = gamma[:, No:, :No]          # Gamma_ph[F, a, i]
    hp = gamma[:, :No, No:]          # Gamma_hp[F, j, b]
    if U is not None:
        ph = np.einsum("Fai,GF->Gai", ph, U)
        hp = np.einsum("Fjb,GF->Gjb", hp, U)
    inv = 1.0 / np.sqrt(np.asarray(V, dtype=np.complex128))
    cTC = np.conj(ph).transpose(0, 2, 1) * inv[:, None, None]   # [G, i, a]
    C = hp * inv[:, None, None]                                 # [G, j, b]
    tau = T2 + np.einsum("ai,bj->abij", T1, T1)
    S = 2.0 * np.einsum("Gia,Gjb,abij->G", cTC, C, tau, optimize=True)
    S -= np.einsum("Gja,Gib,abij->G", cTC, C, tau, optimize=True)
    return S.real


# Catmull-Rom weights of the four points around t in [0, 1)
def _catmull_rom(t):
    t = np.asarray(t, dtype=np.float64)
    t2, t3 = t * t, t * t * t
    return np.stack([0.5 * (-t + 2.0 * t2 - t3),
                     0.5 * (2.0 - 5.0 * t2 + 3.0 * t3),
                     0.5 * (t + 4.0 * t2 - 3.0 * t3),
                     0.5 * (-t2 + t3)], axis=-1)


def tricubic(F, origin, x):
    """Clamped Catmull-Rom tricubic value of the box F[z, y, x] (unit
    spacing, first point at `origin`) at the points x [..., 3] (x, y, z):
    the cell index is the truncated coordinate clamped into the box, the
    four neighbours are clamped to the box, the fraction is not
    (TricubicInterpolation.hpp:231-353)."""
    F = np.asarray(F, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    bounds = np.array([F.shape[2], F.shape[1], F.shape[0]])
    idx = x - np.asarray(origin, dtype=np.float64)
    cell = np.trunc(idx).astype(np.int64)
    cell = np.clip(cell, 0, bounds - 1)
    w = [_catmull_rom(idx[..., d] - cell[..., d]) for d in range(3)]
    out = np.zeros(x.shape[:-1])
    for c in range(4):
        zc = np.clip(cell[..., 2] - 1 + c, 0, bounds[2] - 1)
        for b in range(4):
            yc = np.clip(cell[..., 1] - 1 + b, 0, bounds[1] - 1)
            for a in range(4):
                xc = np.clip(cell[..., 0] - 1 + a, 0, bounds[0] - 1)
                out += w[0][..., a] * w[1][..., b] * w[2][..., c] * F[zc, yc, xc]
    return out


def finite_size_correction(grid, B, S, N=20, vertex_unit=1.0, grid_unit=1.0,
                           potential_unit=1.0):
    """FiniteSizeCorrection::interpolation (FiniteSizeCorrection.cxx:136-311):
    grid [NG, 3] Cartesian momenta, B [3, 3] reciprocal lattice vectors (rows
    Gi, Gj, Gk), S [NG] the structure factor.  Returns (corrected,
    uncorrected); the correction is their difference."""
    grid = np.asarray(grid, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    omega = abs(np.dot(np.cross(B[0], B[1]), B[2]))
    A = np.array([np.cross(B[1], B[2]), np.cross(B[2], B[0]), np.cross(B[0], B[1])]) / omega
    to_units = vertex_unit ** 2 / grid_unit ** 3 / potential_unit
    factor = grid_unit / vertex_unit ** 2 * omega / 2.0 / math.pi / math.pi
    direct = grid @ A.T                              # [NG, 3]
    dmin = np.minimum(direct.min(axis=0), 0.0)       # the box always holds 0
    dmax = np.maximum(direct.max(axis=0), 0.0)
    dims = np.floor(dmax - dmin + 1.5).astype(np.int64)
    origin = np.floor(dmin + 0.5)
    box = np.zeros((dims[2], dims[1], dims[0]))
    pos = (np.floor(direct + 0.5) - origin).astype(np.int64)
    box[pos[:, 2], pos[:, 1], pos[:, 0]] = S
    g2 = np.sum(grid * grid, axis=1)
    keep = np.sqrt(g2) >= 1e-8
    uncorrected = float(np.sum(to_units * factor / g2[keep] * S[keep]))
    r = np.arange(-N, N)
    a, b, c = np.meshgrid(r, r, r, indexing="ij")
    shifts = (np.outer(a.ravel(), B[0]) + np.outer(b.ravel(), B[1])
              + np.outer(c.ravel(), B[2])) / (2.0 * N)
    total = 0.0
    for Gi in grid:
        g = shifts + Gi
        gl2 = np.sum(g * g, axis=1)
        ok = np.sqrt(gl2) >= 1e-8
        val = tricubic(box, origin, g[ok] @ A.T)
        total += float(np.sum(to_units * val * factor / gl2[ok]))
    corrected = total / N / N / N / 8.0
    return corrected, uncorrected
