"""Mixed precision (split-bf16 MFMA, FP32 accumulation, FP64 refinement).

Bars: a split-bf16 contraction within 5e-5 relative of the FP64 result
(each operand part carries ~16 significant bits); a mixed residuum within
1e-4 relative of the FP64 one; a mixed-then-FP64 run converging to the FP64
energy within 1 microhartree.
"""

import numpy as np
import pytest

from cc4s_amd import api
from cc4s_amd import synthetic as syn
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

CASES = [("ac", "cbij", "abij"), ("Gai", "Gbj", "abij"), ("akic", "cbkj", "abij"),
         ("bkci", "ackj", "abij"), ("klcd", "adkl", "ac"), ("cxdy", "cdij", "xyij"),
         ("Gxc", "Gyd", "cxdy")]
LENS = {"a": 21, "b": 21, "c": 21, "d": 21, "i": 6, "j": 6, "k": 6, "l": 6,
        "G": 77, "x": 9, "y": 7}


@pytest.mark.parametrize("case", CASES, ids=["%s,%s->%s" % c for c in CASES])
def test_mixed_contract_accuracy(case):
    ia, ib, ic = case
    rng = np.random.default_rng(7)
    mk = lambda idx: (rng.standard_normal(tuple(LENS[c] for c in idx))
                      + 1j * rng.standard_normal(tuple(LENS[c] for c in idx)))
    A, B, C = mk(ia), mk(ib), mk(ic)
    for alpha, beta in ((1.0, 0.0), (-0.5, 1.0)):
        ref = alpha * np.einsum(f"{ia},{ib}->{ic}", A, B) + beta * C
        out = api.contract(alpha, A, ia, B, ib, beta, C, ic,
                           precision=api.PRECISION_MIXED)
        assert rel_err(out, ref) < 5e-5, (case, alpha, beta, rel_err(out, ref))


@pytest.mark.parametrize("terms", ["12", "9"])
@pytest.mark.parametrize("case", CASES, ids=["%s,%s->%s" % c for c in CASES])
def test_half_split_contract_accuracy(case, terms, monkeypatch):
    """Half-precision split parts with power-of-two operand scaling: about
    22 significant bits per part, the FP32 sums bound the error; operands
    spanning many decades (scaled by 1e-6) keep it.  With 9 products per
    complex multiply-add (Re Re, Im Im and (Re + Im)(Re + Im), the scale
    taken from max |Re| + |Im|) the same bar holds."""
    monkeypatch.setenv("CC4S_MIXED_SPLIT", "half")
    monkeypatch.setenv("CC4S_MIXED_TERMS", terms)
    ia, ib, ic = case
    rng = np.random.default_rng(7)
    mk = lambda idx: (rng.standard_normal(tuple(LENS[c] for c in idx))
                      + 1j * rng.standard_normal(tuple(LENS[c] for c in idx)))
    A, B, C = mk(ia), mk(ib), mk(ic)
    A *= 1e-6 * np.exp(rng.uniform(-8, 0, A.shape))
    for alpha, beta in ((1.0, 0.0), (-0.5, 1.0)):
        ref = alpha * np.einsum(f"{ia},{ib}->{ic}", A, B) + beta * C
        out = api.contract(alpha, A, ia, B, ib, beta, C, ic,
                           precision=api.PRECISION_MIXED)
        assert rel_err(out, ref) < 2e-6, (case, alpha, beta, rel_err(out, ref))


@pytest.mark.parametrize("split,terms", [("bf16", "12"), ("half", "9")])
def test_mixed_contract_conjugation_and_large_k(split, terms, monkeypatch):
    monkeypatch.setenv("CC4S_MIXED_SPLIT", split)
    monkeypatch.setenv("CC4S_MIXED_TERMS", terms)
    rng = np.random.default_rng(3)
    M, N, K = 300, 200, 9000
    A = rng.standard_normal((M, K)) + 1j * rng.standard_normal((M, K))
    B = rng.standard_normal((K, N)) + 1j * rng.standard_normal((K, N))
    C = np.zeros((M, N), complex)
    out = api.contract(1.0, A, "mk", B, "kn", 0.0, C, "mn", conj_a=True,
                       precision=api.PRECISION_MIXED)
    assert rel_err(out, np.conj(A) @ B) < 5e-5


def test_mixed_rejects_real_field():
    with pytest.raises(api.CcsdError, match="complex"):
        api.contract(1.0, np.ones((3, 4)), "mk", np.ones((4, 2)), "kn", 0.0,
                     np.zeros((3, 2)), "mn", precision=api.PRECISION_MIXED)


@pytest.mark.parametrize("split,terms,bound", [("bf16", "12", 1e-4), ("half", "12", 1e-5),
                                               ("half", "9", 1e-5)])
def test_mixed_residuum_accuracy(split, terms, bound, monkeypatch):
    """At No=12, Nv=60, NG=300 the ring, Gamma and ladder contractions are
    above the mixed threshold (1e9 flops) and run split."""
    monkeypatch.setenv("CC4S_MIXED_SPLIT", split)
    monkeypatch.setenv("CC4S_MIXED_TERMS", terms)
    No, Nv, NG = 12, 60, 300
    inp = syn.generate(No, Nv, NG, "complex", 5)
    T = syn.random_amplitudes(No, Nv, "complex", 9, t1_scale=0.02, t2_scale=0.02)
    out = {}
    for prec in (api.PRECISION_FP64, api.PRECISION_MIXED):
        with api.CoupledCluster(No, Nv, NG, precision=prec) as cc:
            cc.set_inputs(inp.gamma, inp.eps)
            out[prec] = cc.residuum(T)
    e1, e2 = rel_err(out[1][0], out[0][0]), rel_err(out[1][1], out[0][1])
    print(f"split {split} ({terms} products): residuum rel. error R1 {e1:.2e} R2 {e2:.2e}",
          flush=True)
    assert e2 < bound and e1 < bound
    assert e2 > 0.0  # the mixed path ran


@pytest.mark.parametrize("switch", [1e-5, 1e-7, 1e-14],
                         ids=["written_1e-5", "tuned_1e-7", "floor_guard"])
@pytest.mark.parametrize("shape", [(6, 24, 60), (12, 60, 300)])
def test_mixed_run_reaches_fp64_energy(shape, switch):
    """The mixed run converges to the FP64 energy; once switched it stays
    FP64; a switch below the split residuum's rounding floor (1e-14) is
    ended by the stall guard instead of running to maxIterations."""
    No, Nv, NG = shape
    print(f"mixed run {shape} switch {switch}", flush=True)
    inp = syn.generate(No, Nv, NG, "complex", 42, scale_rule="mp2_per_occupied")
    kw = dict(maxIterations=60, energyConvergence=1e-10,
              amplitudesConvergence=1e-8, maxResidua=8)
    res = {}
    for prec in (api.PRECISION_FP64, api.PRECISION_MIXED):
        # the plain switch: FP64 passes right after the split passes
        with api.CoupledCluster(No, Nv, NG, precision=prec, mixedSwitch=switch,
                                mixedRefine=0, **kw) as cc:
            cc.set_inputs(inp.gamma, inp.eps)
            res[prec] = cc.run()
    fp64, mixed = res[0], res[1]
    assert fp64["convergenceReached"] and mixed["convergenceReached"]
    assert abs(mixed["energy"]["correlation"] - fp64["energy"]["correlation"]) < 1e-8
    used = [t["precision_used"] for t in mixed["trace"]]
    for t in mixed["trace"]:
        print(f"{t['iteration']:3d} {'mixed' if t['precision_used'] else 'fp64 '} "
              f"E={t['energy']:.13f} dR={t['residuum_norm']:.3e}", flush=True)
    assert used[0] == api.PRECISION_MIXED and used[-1] == api.PRECISION_FP64
    # once switched, FP64 to the end; only FP64 passes declare convergence
    first64 = used.index(api.PRECISION_FP64)
    assert all(u == api.PRECISION_FP64 for u in used[first64:])
    assert not any(t["converged"] for t in mixed["trace"][:first64])
    # the stall guard costs the passes spent at the rounding floor
    assert mixed["iterations"] <= fp64["iterations"] + (12 if switch < 1e-12 else 4)


@pytest.mark.parametrize("placement", ["device", "host", "host_all"])
def test_restart_keeping_recent_pairs(placement, monkeypatch):
    """A DIIS restart at the switch that keeps the two most recent pairs
    (moved to the first slots, wherever the slots live) still converges to
    the FP64 energy."""
    print(f"restart keep 2, DIIS slots {placement}", flush=True)
    monkeypatch.setenv("CC4S_DIIS_PLACEMENT", placement)
    monkeypatch.setenv("CC4S_DIIS_CHUNK_BYTES", "1000000")
    No, Nv, NG = 12, 60, 300
    inp = syn.generate(No, Nv, NG, "complex", 42, scale_rule="mp2_per_occupied")
    kw = dict(maxIterations=60, energyConvergence=1e-10, amplitudesConvergence=1e-8,
              maxResidua=8)
    with api.CoupledCluster(No, Nv, NG, **kw) as cc:
        cc.set_inputs(inp.gamma, inp.eps)
        fp64 = cc.run()
    with api.CoupledCluster(No, Nv, NG, precision=api.PRECISION_MIXED, mixedSwitch=1e-6,
                            mixedRestart=True, mixedRestartKeep=2, mixedRefine=0,
                            **kw) as cc:
        cc.set_inputs(inp.gamma, inp.eps)
        mixed = cc.run()
    assert mixed["convergenceReached"]
    assert abs(mixed["energy"]["correlation"] - fp64["energy"]["correlation"]) < 1e-8


@pytest.mark.parametrize("refine", [1, 2, 3])
def test_defect_corrected_refinement(refine):
    """After the split passes, one FP64 anchor pass measures the defect of
    the split residuum; split passes corrected by it continue, then FP64
    passes finish.  The run reaches the FP64 energy, the passes come in that
    order, and only FP64 passes (anchor or plain) declare convergence; with
    mode 3 a corrected pass may end the run."""
    No, Nv, NG = 12, 60, 300
    print(f"defect-corrected refinement, mode {refine}", flush=True)
    inp = syn.generate(No, Nv, NG, "complex", 42, scale_rule="mp2_per_occupied")
    kw = dict(maxIterations=60, energyConvergence=1e-10, amplitudesConvergence=1e-8,
              maxResidua=8)
    with api.CoupledCluster(No, Nv, NG, **kw) as cc:
        cc.set_inputs(inp.gamma, inp.eps)
        fp64 = cc.run()
    with api.CoupledCluster(No, Nv, NG, precision=api.PRECISION_MIXED, mixedSwitch=1e-6,
                            mixedRefine=refine, mixedSplit="half", mixedMaxK=2048,
                            **kw) as cc:
        cc.set_inputs(inp.gamma, inp.eps)
        mixed = cc.run()
    kinds = "".join("FSAC"[t["pass_kind"]] for t in mixed["trace"])
    for t in mixed["trace"]:
        print(f"{t['iteration']:3d} {'FSAC'[t['pass_kind']]} E={t['energy']:.13f} "
              f"dR={t['residuum_norm']:.3e}", flush=True)
    assert mixed["convergenceReached"]
    assert abs(mixed["energy"]["correlation"] - fp64["energy"]["correlation"]) < 1e-8
    import re
    assert re.fullmatch(r"S+AC*F*", kinds), kinds
    declaring = (0, 2, 3) if refine == 3 else (0, 2)
    assert all(t["pass_kind"] in declaring for t in mixed["trace"] if t["converged"])
    if refine == 3:
        assert mixed["trace"][-1]["converged"]


@pytest.mark.parametrize("opts", [dict(method="Drccd"), dict(levelShift=0.3)],
                         ids=["drccd", "levelShift"])
def test_refinement_with_other_methods_and_shifts(opts):
    """The anchored refinement (default) with method Drccd and with a level
    shift: the mixed run reaches the FP64 run's energy."""
    No, Nv, NG = 12, 60, 300
    print(f"mixed refinement {opts}", flush=True)
    inp = syn.generate(No, Nv, NG, "complex", 42, scale_rule="mp2_per_occupied")
    kw = dict(maxIterations=80, energyConvergence=1e-10, amplitudesConvergence=1e-8,
              maxResidua=8, **opts)
    res = {}
    for prec in (api.PRECISION_FP64, api.PRECISION_MIXED):
        with api.CoupledCluster(No, Nv, NG, precision=prec, **kw) as cc:
            cc.set_inputs(inp.gamma, inp.eps)
            res[prec] = cc.run()
    assert res[0]["convergenceReached"] and res[1]["convergenceReached"]
    assert abs(res[1]["energy"]["correlation"] - res[0]["energy"]["correlation"]) < 1e-8
    assert all(t["pass_kind"] in (0, 2) for t in res[1]["trace"] if t["converged"])
