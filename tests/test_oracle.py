"""Pins the FP64 oracle (CPU only).

Each check is chosen so that a plausible mistake in the restatement of
Ccsd.cxx / CoupledCluster.cxx / DiisMixer.cxx fails it: the Gamma-factorised
residuum must equal the independent V-slice formulation of CcsdReference.cxx,
the first iteration must give the MP2 closed form, DIIS with one vector must
be Jacobi, the fixed point must not depend on levelShift or the ladder slice
size, and V slices must equal brute-force sums over G.
"""

import hashlib

import numpy as np
import pytest

from cc4s_amd import synthetic as syn
from oracle import ccsd_oracle as O
from oracle.ccsd_vbased import residuum_vbased
from tests.conftest import rel_err


def _problem(No, Nv, NG, field, seed, slice_size=None, ppl=True):
    inp = syn.generate(No, Nv, NG, field, seed)
    return O.CcsdProblem(inp.gamma, inp.eps, No, field, ppl=ppl,
                         integrals_slice_size=slice_size), inp


def _amps(No, Nv, field, seed, scale=0.1):
    T1, T2 = syn.random_amplitudes(No, Nv, field, seed, t1_scale=scale,
                                   t2_scale=scale)
    if field == "real":
        return T1.real.copy(), T2.real.copy()
    return T1, T2


@pytest.mark.parametrize("field", ["complex", "real"])
@pytest.mark.parametrize("shape", [(2, 5, 4), (3, 7, 9), (4, 9, 6)])
def test_factorised_equals_vbased(field, shape):
    No, Nv, NG = shape
    P, _ = _problem(No, Nv, NG, field, seed=11, slice_size=2)
    T = _amps(No, Nv, field, 5)
    R1, R2 = P.residuum(T)
    Q1, Q2 = residuum_vbased(P.slices, field, T, No, Nv)
    assert rel_err(R1, Q1) < 1e-12
    assert rel_err(R2, Q2) < 1e-12


def test_real_path_complex_symmetric_vertex():
    # halfGrid vertices of real orbitals: Gamma^G_pq = Gamma^G_qp with Im != 0
    No, Nv, NG = 3, 7, 5
    rng = np.random.default_rng(1)
    g = rng.standard_normal((NG, No + Nv, No + Nv)) + 1j * rng.standard_normal(
        (NG, No + Nv, No + Nv))
    g = 0.1 * (g + g.transpose(0, 2, 1))
    P = O.CcsdProblem(g, syn.eigen_energies(No, Nv, 1), No, "real",
                      integrals_slice_size=2)
    T = _amps(No, Nv, "real", 9)
    R1, R2 = P.residuum(T)
    Q1, Q2 = residuum_vbased(P.slices, "real", T, No, Nv)
    assert rel_err(R1, Q1) < 1e-12
    assert rel_err(R2, Q2) < 1e-12


def test_mutation_is_detected():
    """A dropped term in the factorised residuum is visible to the check."""
    No, Nv, NG = 3, 7, 5
    P, _ = _problem(No, Nv, NG, "complex", seed=11)
    T = _amps(No, Nv, "complex", 5)
    R1, R2 = P.residuum(T)
    Q1, Q2 = residuum_vbased(P.slices, "complex", T, No, Nv)
    # Xklij T2 term of Ccsd.cxx:543 removed
    Xklij = O.ein("klcd,cdij->klij", P.V["hhpp"],
                  T[1] + O.ein("ai,bj->abij", T[0], T[0]))
    R2_mut = R2 - O.ein("klij,abkl->abij", Xklij, T[1])
    assert rel_err(R2_mut, Q2) > 1e-4


@pytest.mark.parametrize("field", ["complex", "real"])
def test_first_iteration_is_mp2(field):
    P, _ = _problem(4, 12, 30, field, seed=3)
    r = O.coupled_cluster(P, max_iterations=1)
    assert abs(r["energy"]["secondOrder"] - O.mp2_energy(P)) < 1e-14
    # first residuum: R1 = 0, R2 = V (Ccsd.cxx:97-102, 404-409)
    R1, R2 = P.residuum(None)
    assert np.all(R1 == 0) and np.array_equal(R2, P.V["pphh"])


@pytest.mark.parametrize("field", ["complex", "real"])
def test_vertex_integrals_brute_force(field):
    No, Nv, NG = 2, 3, 4
    inp = syn.generate(No, Nv, NG, field, seed=2)
    S = O.slice_vertex(inp.gamma, No)
    V = O.vertex_integrals(S, field, O.COMPLEX_SLICES)
    g = inp.gamma
    rng = {"h": range(0, No), "p": range(No, No + Nv)}
    for name, v in V.items():
        LO, RO, LI, RI = name
        for ip, p in enumerate(rng[LO]):
            for iq, q in enumerate(rng[RO]):
                for is_, s in enumerate(rng[LI]):
                    for ir, r in enumerate(rng[RI]):
                        if field == "complex":
                            ref = sum(np.conj(g[G, s, p]) * g[G, q, r] for G in range(NG))
                        else:
                            ref = sum(g[G, p, s].real * g[G, q, r].real
                                      + g[G, p, s].imag * g[G, q, r].imag
                                      for G in range(NG))
                        assert abs(v[ip, iq, is_, ir] - ref) < 1e-14


def test_diis_single_vector_is_jacobi():
    P, _ = _problem(3, 8, 12, "complex", seed=4)
    mixer = O.DiisMixer(1, complex_field=True)
    A = (np.ones((8, 3), complex), np.ones((8, 8, 3, 3), complex))
    R = (0.1 * A[0], 0.1 * A[1])
    mixer.append(A, R)
    assert abs(mixer.last_coefficients[0] - 1.0) < 1e-15
    # full loop with one DIIS vector equals the plain Jacobi update
    rj = O.coupled_cluster(P, max_iterations=3, max_residua=1)
    T = None
    for _ in range(3):
        T = O.residuum_to_amplitudes(P, P.residuum(T), T)
    e, _, _ = O.energy(P, T)
    assert abs(rj["energy"]["correlation"] - e.real) < 1e-15


def test_diis_coefficients_sum_to_one():
    mixer = O.DiisMixer(4, complex_field=False)
    rng = np.random.default_rng(0)
    for _ in range(6):
        A = (rng.standard_normal((5, 2)), rng.standard_normal((5, 5, 2, 2)))
        R = (rng.standard_normal((5, 2)), rng.standard_normal((5, 5, 2, 2)))
        mixer.append(A, R)
        assert abs(sum(mixer.last_coefficients) - 1.0) < 1e-12


@pytest.mark.parametrize("field", ["complex", "real"])
def test_fixed_point_independent_of_path(field):
    """levelShift and integralsSliceSize change the path, not the result
    (CoupledCluster.cxx:264-280, Ccsd.cxx:287-332)."""
    kw = dict(max_iterations=60, energy_convergence=1e-12,
              amplitudes_convergence=1e-10, max_residua=6)
    P1, _ = _problem(4, 12, 20, field, seed=8, slice_size=3)
    P2, _ = _problem(4, 12, 20, field, seed=8, slice_size=12)
    r1 = O.coupled_cluster(P1, **kw)
    r2 = O.coupled_cluster(P2, level_shift=0.3, **kw)
    assert r1["convergenceReached"] and r2["convergenceReached"]
    assert abs(r1["energy"]["correlation"] - r2["energy"]["correlation"]) < 1e-10
    # at the fixed point one more Jacobi step leaves T unchanged
    T = r1["amplitudes"]
    Tn = O.residuum_to_amplitudes(P1, P1.residuum(T), T)
    assert rel_err(Tn[1], T[1]) < 1e-7


def test_generator_is_deterministic_and_seeded():
    a = syn.generate(3, 6, 10, "complex", seed=5)
    b = syn.generate(3, 6, 10, "complex", seed=5)
    c = syn.generate(3, 6, 10, "complex", seed=6)
    assert np.array_equal(a.gamma, b.gamma) and np.array_equal(a.eps, b.eps)
    assert not np.array_equal(a.gamma, c.gamma)
    # Hermitian in (p, q), max |V_abij| = 0.05, sorted energy blocks
    g = a.gamma
    assert np.allclose(g, np.conj(g.transpose(0, 2, 1)), atol=0, rtol=0)
    assert abs(syn.max_abs_vpphh(g, 3, "complex") - 0.05) < 1e-15
    assert np.all(np.diff(a.eps[:3]) >= 0) and np.all(np.diff(a.eps[3:]) >= 0)
    assert a.eps[2] < syn.FERMI_ENERGY < a.eps[3]
    r = syn.generate(3, 6, 10, "real", seed=5)
    assert np.all(r.gamma.imag == 0)
    assert np.array_equal(r.gamma, r.gamma.transpose(0, 2, 1))


def test_define_holes_and_particles():
    eps = np.array([-1.0, -0.5, 0.3, 1.0])
    No, eh, ep = O.define_holes_and_particles(eps, -0.1)
    assert No == 2 and list(eh) == [-1.0, -0.5] and list(ep) == [0.3, 1.0]
    with pytest.raises(ValueError, match="below all"):
        O.define_holes_and_particles(eps, -2.0)
    with pytest.raises(ValueError, match="above all"):
        O.define_holes_and_particles(eps, 2.0)


def test_golden_real_workload_generator_bytes():
    """SHA-256 of the real No=8, Nv=40, NG=200 input bytes (cc4s layout) is fixed."""
    cfg = syn.WORKLOADS["real_no8_nv40_ng200"]
    inp = syn.generate_config(cfg)
    import json, pathlib
    gold = json.loads((pathlib.Path(__file__).parent / "golden" /
                       "real_no8_nv40_ng200.json").read_text())
    h = hashlib.sha256(np.asfortranarray(inp.gamma).tobytes(order="F")).hexdigest()
    he = hashlib.sha256(inp.eps.tobytes()).hexdigest()
    assert h == gold["gamma_sha256"]
    assert he == gold["eps_sha256"]


def test_ueg_known_answer_from_the_reference_test_suite():
    """The reference's own recorded energies: restating UegVertexGenerator.cxx
    for rs=1, No=7, Nv=26 reproduces the inputs of
    test/tests/ueg/rs1.0-7occ-26virt, and the oracle's MP2 and CCSD energies
    equal correct.out.yaml:125-157 (the reference's test tolerance is 1e-7)."""
    from oracle import ueg
    gamma, eps, fermi = ueg.generate(7, 26, 1.0)
    assert gamma.shape == (257, 33, 33)
    No = O.define_holes_and_particles(eps, fermi)[0]
    assert No == 7
    ref = ueg.UEG_RS1_NO7_NV26
    s = ref["settings"]
    P = O.CcsdProblem(gamma, eps, No, "complex",
                      integrals_slice_size=s["integralsSliceSize"])
    assert abs(O.mp2_energy(P) - ref["Ccsd"]["secondOrder"]) < 1e-12
    r = O.coupled_cluster(P, max_iterations=s["maxIterations"],
                          energy_convergence=s["energyConvergence"],
                          amplitudes_convergence=s["amplitudesConvergence"],
                          max_residua=s["maxResidua"])
    assert r["convergenceReached"]
    for key in ("correlation", "direct", "exchange", "secondOrder"):
        assert abs(r["energy"][key] - ref["Ccsd"][key]) < 1e-10, key
    # (T) of the converged CCSD amplitudes (PerturbativeTriplesReference)
    e_t = O.perturbative_triples(gamma, eps, No, "complex", r["amplitudes"])
    assert abs(e_t - ref["Triples"]["correlation"]) < 1e-12
    # Drccd, the step before Ccsd in the same operation file
    P = O.CcsdProblem(gamma, eps, No, "complex", method="Drccd")
    r = O.coupled_cluster(P, max_iterations=s["maxIterations"],
                          energy_convergence=s["energyConvergence"],
                          amplitudes_convergence=s["amplitudesConvergence"],
                          max_residua=s["maxResidua"])
    assert r["convergenceReached"]
    for key in ("correlation", "direct", "exchange", "secondOrder"):
        assert abs(r["energy"][key] - ref["Drccd"][key]) < 1e-10, key


def test_second_order_singles_from_the_one_body_hamiltonian():
    """The singles term of SecondOrderPerturbationTheory is the lambda^2
    coefficient of the exact ground-state energy of the one-body Hamiltonian
    diag(e) + lambda (f_ph + f_ph^dagger) (g times the sum of its No lowest
    eigenvalues), taken here by Richardson-extrapolated finite differences."""
    No, Nv, NG = 3, 6, 8
    inp = syn.generate(No, Nv, NG, "complex", seed=21)
    P = O.CcsdProblem(inp.gamma, inp.eps, No, "complex")
    rng = np.random.default_rng(5)
    f = 0.05 * (rng.standard_normal((Nv, No)) + 1j * rng.standard_normal((Nv, No)))
    Np = No + Nv

    def ground(lam, g=2.0):
        h = np.diag(np.asarray(inp.eps, dtype=complex))
        h[No:, :No] += lam * f
        h[:No, No:] += lam * np.conj(f).T
        return g * np.sum(np.linalg.eigvalsh(h)[:No])

    assert Np == inp.gamma.shape[1]
    e0 = ground(0.0)
    d = lambda lam: (ground(lam) - e0) / lam ** 2
    lam = 1e-2
    e2 = (4.0 * d(lam / 2) - d(lam)) / 3.0
    s = O.second_order(P, fph=f)["singles"]
    assert abs(s - e2) < 1e-6 * abs(e2)
    assert O.second_order(P, fph=f, degeneracy=1.0)["singles"] == pytest.approx(s / 2, rel=1e-15)


def _spin_orbitals(gamma, eps, No):
    """The same closed-shell system in spin orbitals: holes (up, down), then
    particles (up, down); the vertex is diagonal in spin."""
    NG, Np = gamma.shape[0], gamma.shape[1]
    Nv = Np - No
    up = list(range(No)) + [None] * No + list(range(No, Np)) + [None] * Nv
    dn = [None] * No + list(range(No)) + [None] * Nv + list(range(No, Np))
    g = np.zeros((NG, 2 * Np, 2 * Np), dtype=gamma.dtype)
    for m in (up, dn):
        idx = [k for k, p in enumerate(m) if p is not None]
        src = [m[k] for k in idx]
        g[np.ix_(range(NG), idx, idx)] = gamma[np.ix_(range(NG), src, src)]
    e = np.concatenate([eps[:No], eps[:No], eps[No:], eps[No:]])
    return g, e


def test_second_order_degeneracy_by_spin_orbitals():
    """Degeneracy 1 (open shell) on the spin-orbital restatement of a
    closed-shell system gives the closed-shell (degeneracy 2) energy: g^2 on
    the direct term and g on the exchange term are the only consistent
    choice, with and without Fock singles."""
    No, Nv, NG = 2, 4, 6
    inp = syn.generate(No, Nv, NG, "complex", seed=23)
    P = O.CcsdProblem(inp.gamma, inp.eps, No, "complex")
    g, e = _spin_orbitals(inp.gamma, np.asarray(inp.eps), No)
    Pso = O.CcsdProblem(g, e, 2 * No, "complex")
    rng = np.random.default_rng(8)
    f = 0.1 * (rng.standard_normal((Nv, No)) + 1j * rng.standard_normal((Nv, No)))
    fso = np.zeros((2 * Nv, 2 * No), dtype=complex)
    fso[:Nv, :No] = f
    fso[Nv:, No:] = f
    for fc, fs in ((None, None), (f, fso)):
        closed = O.second_order(P, fph=fc)
        so = O.second_order(Pso, fph=fs, degeneracy=1.0)
        assert abs(so["correlation"] - closed["correlation"]) < 1e-13 * abs(closed["correlation"])
        assert abs(so["singles"] - closed["singles"]) < 1e-15 + 1e-13 * abs(closed["singles"])


@pytest.mark.parametrize("field", ["complex", "real"])
def test_mp2_pair_energies_and_triples_star(field):
    """The finite-basis MP2 pair energies of PerturbativeTriplesStar.hpp sum to
    the second-order energy of SecondOrderPerturbationTheory (another
    denominator sign convention and the exchange read as V_abji), so given
    those pair energies the (T*) correction is the (T) energy itself, and it
    scales with the given pair energies."""
    No, Nv, NG = 3, 8, 10
    inp = syn.generate(No, Nv, NG, field, seed=29)
    P = O.CcsdProblem(inp.gamma, inp.eps, No, field)
    e = O.mp2_pair_energies(P)
    assert e.shape == (No, No)
    assert abs(np.sum(e).real - O.mp2_energy(P)) < 1e-14 * abs(O.mp2_energy(P))
    assert abs(np.sum(e).imag) < 1e-14 * abs(O.mp2_energy(P))
    assert O.triples_star(P, -0.0123, e.real) == pytest.approx(-0.0123, rel=1e-14)
    assert O.triples_star(P, -0.0123, 1.25 * e.real) == pytest.approx(1.25 * -0.0123, rel=1e-14)


@pytest.mark.parametrize("seed", [3, 9])
def test_drccd_direct_energy_is_the_rpa_plasmon_formula(seed):
    """Direct ring CCD is the random-phase approximation (Scuseria,
    Henderson and Sorensen, J. Chem. Phys. 129, 231101 (2008)): for real
    orbitals the converged direct energy of the Drccd method equals
    (sum_n w_n - tr A) / 2 with A = D + K, B = K, K_(ai),(bj) = 2 V_abij and
    w_n^2 the eigenvalues of (A - B)^1/2 (A + B) (A - B)^1/2."""
    No, Nv, NG = 3, 8, 10
    inp = syn.generate(No, Nv, NG, "real", seed=seed, scale_rule="mp2_per_occupied")
    P = O.CcsdProblem(inp.gamma, inp.eps, No, "real", method="Drccd")
    r = O.coupled_cluster(P, max_iterations=400, energy_convergence=1e-14,
                          amplitudes_convergence=1e-13, max_residua=8)
    assert r["convergenceReached"]
    n = No * Nv
    K = 2.0 * np.real(P.V["pphh"]).transpose(0, 2, 1, 3).reshape(n, n)
    D = (P.eps_p[:, None] - P.eps_h[None, :]).reshape(n)
    M = np.sqrt(D)[:, None] * (np.diag(D) + 2.0 * K) * np.sqrt(D)[None, :]
    w = np.sqrt(np.linalg.eigvalsh(M))
    e_rpa = 0.5 * (w.sum() - D.sum() - np.trace(K))
    assert abs(r["energy"]["direct"] - e_rpa) < 1e-12
