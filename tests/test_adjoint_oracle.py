"""Pins the oracle's energy-adjoint functions (CPU only).

``vertex_energy`` is checked against the energy through the V slices
(CoupledCluster.cxx:194-253 on VertexCoulombIntegrals.cxx's pphh slice), and
``vertex_gradient`` against central finite differences of ``vertex_energy``
in every real component of the vertex (the energy is quadratic in Gamma, so
the central difference is exact up to rounding), for both fields and for
vertices without the Hermitian symmetry.  The structure factor must sum to
the energy and equal the derivative along a scaling of one G slice.
"""

import numpy as np
import pytest

from cc4s_amd import synthetic as syn
from oracle import ccsd_oracle as O
from tests.conftest import rel_err


def _vertex(NG, Np, field, seed, hermitian):
    rng = np.random.default_rng(seed)
    g = 0.3 * (rng.standard_normal((NG, Np, Np)) + 1j * rng.standard_normal((NG, Np, Np)))
    if hermitian:
        g = 0.5 * (g + np.conj(g.transpose(0, 2, 1)))
    return g


def _amps(No, Nv, field, seed):
    T1, T2 = syn.random_amplitudes(No, Nv, field, seed, t1_scale=0.2, t2_scale=0.2)
    if field == "real":
        return T1.real.copy(), T2.real.copy()
    return T1, T2


@pytest.mark.parametrize("field", ["complex", "real"])
@pytest.mark.parametrize("hermitian", [True, False])
def test_vertex_energy_equals_energy_of_v_slices(field, hermitian):
    No, Nv, NG = 3, 5, 4
    g = _vertex(NG, No + Nv, field, 3, hermitian)
    T = _amps(No, Nv, field, 4)
    P = O.CcsdProblem(g, syn.eigen_energies(No, Nv, 1), No, field)
    E, _, _ = O.energy(P, T)
    assert abs(O.vertex_energy(g, T, No, field) - E.real) < 1e-13 * abs(E)


@pytest.mark.parametrize("field", ["complex", "real"])
@pytest.mark.parametrize("hermitian", [True, False])
def test_gradient_central_differences(field, hermitian):
    No, Nv, NG = 2, 3, 2
    Np = No + Nv
    g = _vertex(NG, Np, field, 7, hermitian)
    T = _amps(No, Nv, field, 8)
    grad, _ = O.vertex_gradient(g, T, No, field)
    h = 1e-5
    fd = np.zeros_like(grad)
    for G in range(NG):
        for p in range(Np):
            for q in range(Np):
                for unit in (1.0, 1j):
                    gp = g.copy()
                    gm = g.copy()
                    gp[G, p, q] += h * unit
                    gm[G, p, q] -= h * unit
                    d = (O.vertex_energy(gp, T, No, field)
                         - O.vertex_energy(gm, T, No, field)) / (2 * h)
                    fd[G, p, q] += d * unit
    assert rel_err(grad, fd) < 1e-8
    # the hh and pp blocks do not enter the energy
    assert np.max(np.abs(fd[:, :No, :No])) < 1e-9
    assert np.max(np.abs(fd[:, No:, No:])) < 1e-9


@pytest.mark.parametrize("field", ["complex", "real"])
def test_structure_factor(field):
    No, Nv, NG = 3, 6, 5
    g = _vertex(NG, No + Nv, field, 12, True)
    T = _amps(No, Nv, field, 13)
    E = O.vertex_energy(g, T, No, field)
    _, S = O.vertex_gradient(g, T, No, field)
    assert abs(np.sum(S) - E) < 1e-12 * abs(E)
    h = 1e-5
    for G in range(NG):
        gp = g.copy()
        gm = g.copy()
        gp[G] *= 1 + h
        gm[G] *= 1 - h
        # E(G slice scaled by 1 + t) = E_rest + (1 + t)^2 S(G)
        d = (O.vertex_energy(gp, T, No, field) - O.vertex_energy(gm, T, No, field)) / (4 * h)
        assert abs(d - S[G]) < 1e-8 * np.max(np.abs(S))
