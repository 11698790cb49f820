"""The fused particle-particle ladder (one kernel forms the V tiles of
Ccsd.cxx:579-600 in LDS and contracts them with tau; no V slab in HBM)
against the FP64 oracle and against the slab path, through the C ABI.

Bar: one residuum within 1e-12 relative of the oracle (and of the slab
path); a converged run within 1e-9 Eh and 1e-8 relative of the oracle's,
with the same iteration count.  The shapes cover every edge of the kernel's
tiling: pair blocks cut by the slice size (x or y counts below 4), Nv not a
multiple of the 16-wide (c, d) chunk, No^2 below one 16-column block, No^2
between the two column-block widths, No^2 above 1024 (two column groups),
a slice size above Nv, a vertex that is not Hermitian (separate right-hand
dressed vertex) and conjugated operand views.
"""

import json
import pathlib

import numpy as np
import pytest

from cc4s_amd import api
from cc4s_amd import synthetic as syn
from oracle import ccsd_oracle as O
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).parent / "golden"
R_TOL = 1e-12
E_TOL = 1e-9
T_TOL = 1e-8


def _amps(No, Nv, seed, scale=0.05):
    return syn.random_amplitudes(No, Nv, "complex", seed, t1_scale=scale, t2_scale=scale)


@pytest.mark.parametrize("shape", [
    (3, 7, 5, 3), (4, 12, 30, 5), (5, 21, 37, 8), (6, 33, 41, 0), (17, 20, 23, 0),
    (33, 18, 9, 7), (4, 10, 12, 64)],
    ids=["o3v7G5s3", "o4v12G30s5", "o5v21G37s8", "o6v33G41", "o17v20G23",
         "o33v18G9s7", "o4v10G12s64"])
def test_fused_residuum_parity(shape):
    No, Nv, NG, s = shape
    print(f"fused ladder residuum No={No} Nv={Nv} NG={NG} slice {s}", flush=True)
    inp = syn.generate(No, Nv, NG, "complex", 21)
    T = _amps(No, Nv, 13)
    P = O.CcsdProblem(inp.gamma, inp.eps, No, "complex", integrals_slice_size=s or None)
    Q1, Q2 = P.residuum(T)
    out = {}
    for ladder in ("fused", "slab"):
        with api.CoupledCluster(No, Nv, NG, field="complex", integralsSliceSize=s,
                                ladder=ladder) as cc:
            cc.set_inputs(inp.gamma, inp.eps)
            out[ladder] = cc.residuum(T)
    R1, R2 = out["fused"]
    print(f"  R2 vs oracle {rel_err(R2, Q2):.2e}, vs slab {rel_err(R2, out['slab'][1]):.2e}",
          flush=True)
    assert rel_err(R1, Q1) < R_TOL
    assert rel_err(R2, Q2) < R_TOL
    assert rel_err(R2, out["slab"][1]) < R_TOL


def test_fused_residuum_non_hermitian_vertex():
    """Without Gamma_qp = conj Gamma_pq the right-hand dressed vertex is a
    tensor of its own (Ccsd.cxx:560-561) and the left one a conjugated view."""
    No, Nv, NG = 4, 12, 30
    inp = syn.generate(No, Nv, NG, "complex", 21)
    rng = np.random.default_rng(4)
    g = inp.gamma + 0.003 * (rng.standard_normal(inp.gamma.shape)
                             + 1j * rng.standard_normal(inp.gamma.shape))
    P = O.CcsdProblem(g, inp.eps, No, "complex", integrals_slice_size=5)
    T = _amps(No, Nv, 13)
    with api.CoupledCluster(No, Nv, NG, field="complex", integralsSliceSize=5,
                            ladder="fused") as cc:
        cc.set_inputs(g, inp.eps)
        R1, R2 = cc.residuum(T)
    Q1, Q2 = P.residuum(T)
    assert rel_err(R1, Q1) < R_TOL
    assert rel_err(R2, Q2) < R_TOL


def test_fused_run_parity_small_complex():
    """The small complex fixture run to convergence with the fused ladder."""
    gold = json.loads((GOLDEN / "small_complex.json").read_text())
    c = gold["config"]
    kw = dict(maxIterations=c["max_iterations"], energyConvergence=c["energy_convergence"],
              amplitudesConvergence=c["amplitudes_convergence"], maxResidua=c["max_residua"])
    inp = syn.generate(c["No"], c["Nv"], c["NG"], c["field"], c["seed"])
    with api.CoupledCluster(c["No"], c["Nv"], c["NG"], field=c["field"], ladder="fused",
                            **kw) as cc:
        cc.set_inputs(inp.gamma, inp.eps)
        gpu = cc.run()
        T1, T2 = cc.amplitudes()
    P = O.CcsdProblem(inp.gamma, inp.eps, c["No"], c["field"])
    ref = O.coupled_cluster(P, max_iterations=c["max_iterations"],
                            energy_convergence=c["energy_convergence"],
                            amplitudes_convergence=c["amplitudes_convergence"],
                            max_residua=c["max_residua"])
    assert gpu["iterations"] == ref["iterations"] == gold["iterations"]
    assert gpu["convergenceReached"] == ref["convergenceReached"]
    for key in ("correlation", "direct", "exchange", "secondOrder"):
        assert abs(gpu["energy"][key] - ref["energy"][key]) < E_TOL, key
    assert abs(gpu["energy"]["correlation"] - gold["energy"]["correlation"]) < E_TOL
    assert rel_err(T1, ref["amplitudes"][0]) < T_TOL
    assert rel_err(T2, ref["amplitudes"][1]) < T_TOL


def test_fused_in_mixed_context_keeps_slab_for_split_passes(monkeypatch):
    """A mixed context with the fused ladder: split passes use the slab (every
    complex contraction split here), FP64 passes the fused kernel; the run
    reaches the FP64 oracle's energy within the mixed path's 1e-8 Eh."""
    monkeypatch.setenv("CC4S_MIXED_MIN_FLOPS", "0")
    No, Nv, NG = 6, 24, 60
    inp = syn.generate(No, Nv, NG, "complex", 42)
    kw = dict(maxIterations=40, energyConvergence=1e-10, amplitudesConvergence=1e-8,
              maxResidua=8)
    with api.CoupledCluster(No, Nv, NG, field="complex", precision=api.PRECISION_MIXED,
                            mixedSwitch=1e-3, ladder="fused", **kw) as cc:
        cc.set_inputs(inp.gamma, inp.eps)
        gpu = cc.run()
    P = O.CcsdProblem(inp.gamma, inp.eps, No, "complex")
    ref = O.coupled_cluster(P, max_iterations=40, energy_convergence=1e-10,
                            amplitudes_convergence=1e-8, max_residua=8)
    kinds = {t["pass_kind"] for t in gpu["trace"]}
    assert gpu["convergenceReached"] and 1 in kinds and (0 in kinds or 2 in kinds)
    assert abs(gpu["energy"]["correlation"] - ref["energy"]["correlation"]) < 1e-8
