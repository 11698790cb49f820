"""The complex No=64, Nv=512, NG=3000 and No=64, Nv=640, NG=4000 workloads
at full size on the GPU, checked by what the mathematics fixes (no oracle
run fits these sizes):

* the first pass (R = V_abij, Ccsd.cxx:404-409) reproduces the closed-form
  second-order energy of SecondOrderPerturbationTheory.cxx:108-127;
* the residuum of pair-symmetric amplitudes is pair-symmetric,
  R_abij = R_baji (the P(ia,jb) symmetrisation Ccsd.cxx:520 and the mirrored
  ladder blocks :591-596);
* orbital phase gauge: with Gamma^G_pq -> conj(u_p) Gamma^G_pq u_q and the
  amplitudes transformed alike, the energy is unchanged and the residuum
  transforms as R_abij -> conj(u_a u_b) u_i u_j R_abij (tests/properties.py);
* the split 16-bit residuum lies within the half split's bound of the FP64
  residuum at the same amplitudes.

At these sizes the ladder runs several slab chunks per slice pair and the
No=64, Nv=640 context keeps its DIIS history in host memory: paths that small
shapes do not take.  One context at a time is open; each is closed before
the next.  Amplitudes: the first pass's (MP2) doubles plus seeded singles.

Progress goes straight to the terminal (pytest's capture is bypassed), so a
pass of a minute or two is never silent for long.
"""

import time

import numpy as np
import pytest

from cc4s_amd import api
from cc4s_amd import synthetic as syn
from tests import properties as PR

pytestmark = pytest.mark.gpu

GAUGE_TOL = 1e-12        # relative, residuum and energy under the gauge
SYMMETRY_TOL = 1e-13     # relative, R_abij - R_baji
MP2_TOL = 1e-13          # relative, first-pass energy vs closed form
SPLIT_TOL = 1e-5         # relative, half split vs FP64 residuum


@pytest.fixture
def say(capsys):
    t0 = time.perf_counter()

    def _say(msg):
        with capsys.disabled():
            print(f"  [{time.perf_counter() - t0:7.1f} s] {msg}", flush=True)
    return _say


def _inputs(cfg, say):
    say(f"generating the {cfg.name} vertex (scale rule mp2_per_occupied)")
    inp = syn.generate_config(cfg, scale_rule="mp2_per_occupied",
                              mp2_fn=syn.mp2_on_device)
    say(f"vertex ready, c = {inp.scale:.6e}")
    return inp


def _singles(cfg, seed=31, scale=0.005):
    n1 = cfg.Nv * cfg.No
    z = syn.normals(seed, syn.STREAM_AMPLITUDES, np.arange(2 * n1))
    return np.asfortranarray(((z[:n1] + 1j * z[n1:]) * scale).reshape(
        cfg.Nv, cfg.No, order="F"))


def _first_pass(cc, cfg, inp, say):
    """Pass 1 against the closed-form MP2 energy of the context's V_abij;
    returns (pass report, MP2 doubles)."""
    it = cc.iterate()
    say(f"pass 1: E = {it['energy']:.12f}")
    _, V = cc.residuum(None)
    mp2 = PR.mp2_closed_form(V, inp.eps, cfg.No)
    del V
    say(f"closed-form MP2 {mp2:.12f}, |dE| = {abs(it['energy'] - mp2):.2e}")
    assert abs(it["energy"] - mp2) <= MP2_TOL * abs(mp2)
    _, T2 = cc.amplitudes()
    return it, T2


def test_no64_nv512_ng3000_full_size_properties(say):
    cfg = syn.WORKLOADS["complex_no64_nv512_ng3000"]
    inp = _inputs(cfg, say)
    T1 = _singles(cfg)
    with api.CoupledCluster(cfg.No, cfg.Nv, cfg.NG, maxResidua=cfg.max_residua) as cc:
        cc.set_inputs(inp.gamma, inp.eps)
        say("FP64 context ready")
        _, T2 = _first_pass(cc, cfg, inp, say)
        E = cc.energy((T1, T2))[0]
        say("FP64 residuum ...")
        R1, R2 = cc.residuum((T1, T2))
        say("FP64 residuum done")
    asym = PR.pair_asymmetry(R2)
    say(f"pair asymmetry {asym:.2e}")
    assert asym < SYMMETRY_TOL

    with api.CoupledCluster(cfg.No, cfg.Nv, cfg.NG, precision=api.PRECISION_MIXED,
                            mixedSplit="half") as cc:
        cc.set_inputs(inp.gamma, inp.eps)
        say("split residuum ...")
        S1, S2 = cc.residuum((T1, T2))
        say("split residuum done")
    e1, e2 = PR.rel_diff(S1, R1), PR.rel_diff(S2, R2)
    del S1, S2
    say(f"split vs FP64: R1 {e1:.2e}, R2 {e2:.2e}")
    assert 0.0 < e2 < SPLIT_TOL and e1 < SPLIT_TOL

    u = PR.phases(cfg.Np, seed=2026)
    PR.gauge_vertex(inp.gamma, u)
    phi = PR.amplitude_phases(u, cfg.No)
    PR.gauge_amplitudes(T1, T2, phi)
    with api.CoupledCluster(cfg.No, cfg.Nv, cfg.NG, maxResidua=cfg.max_residua) as cc:
        cc.set_inputs(inp.gamma, inp.eps)
        Eg = cc.energy((T1, T2))[0]
        say("gauged FP64 residuum ...")
        G1, G2 = cc.residuum((T1, T2))
        say("gauged FP64 residuum done")
    dE = abs(Eg - E) / abs(E)
    d1 = float(np.max(np.abs(G1 - phi * R1)) / np.max(np.abs(R1)))
    d2 = PR.gauge_defect(G2, R2, phi)
    say(f"gauge: energy {dE:.2e}, R1 {d1:.2e}, R2 {d2:.2e} (relative)")
    assert dE < GAUGE_TOL and d1 < GAUGE_TOL and d2 < GAUGE_TOL


def test_no64_nv640_ng4000_full_size_passes(say):
    """Two passes of the iteration with the DIIS history in host memory:
    pass 1 is the closed-form MP2, pass 2 evaluates the full residuum and
    the DIIS step; its amplitudes stay pair-symmetric and its energy is the
    energy of its amplitudes."""
    cfg = syn.WORKLOADS["complex_no64_nv640_ng4000"]
    inp = _inputs(cfg, say)
    with api.CoupledCluster(cfg.No, cfg.Nv, cfg.NG, maxResidua=cfg.max_residua,
                            maxIterations=cfg.max_iterations,
                            energyConvergence=cfg.energy_convergence,
                            amplitudesConvergence=cfg.amplitudes_convergence) as cc:
        cc.set_inputs(inp.gamma, inp.eps)
        say("FP64 context ready")
        it1, T2 = _first_pass(cc, cfg, inp, say)
        del T2
        say("pass 2 (full residuum, DIIS) ...")
        it2 = cc.iterate()
        say(f"pass 2: E = {it2['energy']:.12f} dR = {it2['residuum_norm']:.3e} "
            f"{it2['seconds']:.1f} s, DIIS history placement {it2['diis_placement']}")
        T1, T2 = cc.amplitudes()
        E = cc.energy((T1, T2))[0]
    assert it2["diis_placement"] == 2          # every slot in host memory
    assert np.isfinite(it2["residuum_norm"]) and it2["residuum_norm"] > 0
    assert abs(it2["energy"] - E.real) <= 1e-12 * abs(E)
    asym = PR.pair_asymmetry(T2)
    say(f"amplitude pair asymmetry {asym:.2e}")
    assert asym < SYMMETRY_TOL
