"""FiniteSizeCorrection (FiniteSizeCorrection.cxx): the oracle's pins and the
library's host interpolation against the oracle (no GPU needed).

Pins of the oracle: the Catmull-Rom tricubic spline returns the grid values
at grid points and reproduces quadratic functions inside the box; the
structure factor at identity singular vectors decomposes the CCSD
correlation energy over G (sum_G V(G) S(G) = E of the same amplitudes, the
energy computed by the oracle's getEnergy restatement); singular vectors act
as a change of the vertex's G basis.
"""

import numpy as np
import pytest

from cc4s_amd import api
from cc4s_amd import synthetic as syn
from oracle import ccsd_oracle as O
from oracle import finite_size as F


def test_tricubic_interpolates_grid_values_and_quadratics():
    rng = np.random.default_rng(1)
    box = rng.standard_normal((6, 7, 8))          # [z, y, x]
    origin = np.array([-3.0, -2.0, -4.0])
    pts = np.array([[x, y, z] for z in range(6) for y in range(7) for x in range(8)],
                   dtype=float) + origin
    assert np.allclose(F.tricubic(box, origin, pts), box.ravel(), rtol=0, atol=1e-14)
    q = lambda p: (0.3 + 0.5 * p[..., 0] - 0.2 * p[..., 1] + 0.7 * p[..., 2]
                   + 0.11 * p[..., 0] ** 2 - 0.05 * p[..., 1] * p[..., 2]
                   + 0.08 * p[..., 2] ** 2)
    zz, yy, xx = np.meshgrid(np.arange(6), np.arange(7), np.arange(8), indexing="ij")
    grid = np.stack([xx, yy, zz], axis=-1) + origin
    vals = q(grid)
    # interior points: every neighbour of the cell lies inside the box
    inner = rng.uniform([1, 1, 1], [5.99, 4.99, 3.99], size=(200, 3)) + origin
    assert np.allclose(F.tricubic(vals, origin, inner), q(inner), rtol=0, atol=1e-12)


def test_structure_factor_decomposes_the_energy():
    No, Nv, NG = 3, 7, 11
    inp = syn.generate(No, Nv, NG, "complex", 5)
    T1, T2 = syn.random_amplitudes(No, Nv, "complex", 3, t1_scale=0.05, t2_scale=0.05)
    V = np.random.default_rng(2).uniform(0.5, 2.0, NG)
    S = F.transition_structure_factor(inp.gamma, No, T1, T2, V)
    P = O.CcsdProblem(inp.gamma, inp.eps, No, "complex")
    E, _, _ = O.energy(P, (T1, T2))
    assert abs(np.sum(V * S) - E.real) < 1e-13 * max(1.0, abs(E))


def test_singular_vectors_change_the_g_basis():
    No, Nv, NG, NGg = 2, 5, 6, 9
    inp = syn.generate(No, Nv, NG, "complex", 7)
    rng = np.random.default_rng(4)
    U = rng.standard_normal((NGg, NG)) + 1j * rng.standard_normal((NGg, NG))
    V = rng.uniform(0.5, 2.0, NGg)
    T1, T2 = syn.random_amplitudes(No, Nv, "complex", 8, t1_scale=0.05, t2_scale=0.05)
    on_grid = np.einsum("Fpq,GF->Gpq", inp.gamma, U)
    assert np.allclose(F.transition_structure_factor(inp.gamma, No, T1, T2, V, U),
                       F.transition_structure_factor(on_grid, No, T1, T2, V),
                       rtol=1e-12, atol=1e-14)


def _lattice(kind):
    if kind == "cubic":
        return 0.9 * np.eye(3)
    # reciprocal vectors of an fcc cell (a bcc lattice), scaled
    return 0.7 * np.array([[-1.0, 1.0, 1.0], [1.0, -1.0, 1.0], [1.0, 1.0, -1.0]])


def _grid(B, radius):
    r = range(-radius, radius + 1)
    pts = [i * B[0] + j * B[1] + k * B[2] for i in r for j in r for k in r
           if i * i + j * j + k * k <= radius * radius]
    return np.array(pts)


@pytest.mark.parametrize("kind", ["cubic", "bcc"])
def test_library_interpolation_matches_the_oracle(kind):
    B = _lattice(kind)
    grid = _grid(B, 3)
    g2 = np.sum(grid * grid, axis=1)
    S = np.exp(-0.4 * g2) * (1.0 + 0.1 * grid[:, 0])
    units = dict(vertex_unit=0.8, grid_unit=1.3, potential_unit=0.6)
    want = F.finite_size_correction(grid, B, S, N=3, **units)
    got = api.finite_size_correction(grid, B, S, N=3, **units)
    assert abs(got[1] - want[1]) <= 1e-12 * abs(want[1])
    assert abs(got[0] - want[0]) <= 1e-12 * abs(want[0])


def test_library_interpolation_refuses_a_flat_lattice():
    B = np.array([[1.0, 0, 0], [0, 1.0, 0], [1.0, 1.0, 0]])
    with pytest.raises(api.CcsdError, match="linearly dependent"):
        api.finite_size_correction(np.zeros((1, 3)), B, np.ones(1), N=2)


def test_finite_size_integration_properties():
    """Pins of the integration that do not go through the spline: for a
    constant S the interpolant is that constant (Catmull-Rom weights sum to
    one) when the grid fills its bounding box (a parallelepiped of lattice
    points), so `corrected` is the plain lattice sum of c / |G + s|^2 over
    every fine-grid point; the vertex unit cancels; the result is linear in
    S."""
    B = _lattice("bcc")
    r2 = range(-2, 3)
    grid = np.array([i * B[0] + j * B[1] + k * B[2] for i in r2 for j in r2 for k in r2])
    N = 2
    omega = abs(np.dot(np.cross(B[0], B[1]), B[2]))
    c = 0.37
    corrected, uncorrected = F.finite_size_correction(grid, B, np.full(len(grid), c), N=N)
    r = np.arange(-N, N)
    shifts = np.array([(a * B[0] + b * B[1] + cc * B[2]) / (2 * N) for a in r for b in r for cc in r])
    pts = (grid[:, None, :] + shifts[None, :, :]).reshape(-1, 3)
    q2 = np.sum(pts * pts, axis=1)
    lattice_sum = np.sum(1.0 / q2[q2 > 1e-16])
    pref = omega / (2.0 * np.pi ** 2)
    assert abs(corrected - pref * c * lattice_sum / (8 * N ** 3)) < 1e-12 * abs(corrected)
    g2 = np.sum(grid * grid, axis=1)
    assert abs(uncorrected - pref * c * np.sum(1.0 / g2[g2 > 1e-16])) < 1e-12 * abs(uncorrected)
    S = np.cos(grid[:, 0]) + 0.2 * grid[:, 1] ** 2
    a = F.finite_size_correction(grid, B, S, N=N, vertex_unit=0.3, grid_unit=1.1)
    b = F.finite_size_correction(grid, B, S, N=N, vertex_unit=2.0, grid_unit=1.1)
    assert np.allclose(a, b, rtol=1e-13)
    two = F.finite_size_correction(grid, B, 2.0 * S, N=N, grid_unit=1.1)
    one = F.finite_size_correction(grid, B, S, N=N, grid_unit=1.1)
    assert np.allclose(two, 2.0 * np.array(one), rtol=1e-13)


def test_ueg_momentum_grid_matches_the_vertex():
    """The momentum grid and Coulomb potential of the electron-gas vertex
    (UegVertexGenerator.cxx:213-296): every nonzero Gamma^G_pq has
    |Gamma^G_pq|^2 = V(G), and the momenta of the pair it couples differ by
    the grid vector of G, k_p - k_q = G (the vertex itself reproduces the
    reference's recorded energies, test_oracle.py)."""
    from oracle import ueg
    No, Nv = 7, 26
    gamma, eps, fermi = ueg.generate(No, Nv, 1.0)
    grid, B, V = ueg.momentum_grid(No, Nv, 1.0)
    assert gamma.shape[0] == grid.shape[0] == V.shape[0]
    b = B[0, 0]
    Np = No + Nv
    maxG = int((5.0 * Np) ** (1.0 / 3.0))
    import itertools
    k = sorted(itertools.product(range(-maxG, maxG + 1), repeat=3),
               key=lambda g: g[0] ** 2 + g[1] ** 2 + g[2] ** 2)[:Np]
    k = b * np.array(k, dtype=float)
    nz = np.argwhere(np.abs(gamma) > 0)
    assert len(nz) > 0
    for G, p, q in nz:
        assert abs(abs(gamma[G, p, q]) ** 2 - V[G]) <= 1e-14 * V[G]
        assert np.allclose(k[p] - k[q], grid[G])


def test_library_interpolation_refuses_an_empty_fine_grid():
    B = _lattice("cubic")
    grid = _grid(B, 2)
    with pytest.raises(api.CcsdError, match="N > 0"):
        api.finite_size_correction(grid, B, np.ones(len(grid)), N=0)
