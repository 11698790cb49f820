"""The cc4s operation-file driver end to end on the GPU.

A cc4s operation file (Read, DefineHolesAndParticles, SliceOperator,
VertexCoulombIntegrals, CoupledCluster, Write) runs on cc4s-format input files;
the report's energies must match the golden fixtures / the oracle to 1e-9 Eh,
the written Amplitudes the oracle's to 1e-8 relative, and a restart from the
written Amplitudes (initialAmplitudes) must reproduce the energy.
"""

import json
import pathlib
import subprocess

import numpy as np
import pytest
import yaml

from cc4s_amd import synthetic as syn
from oracle import ccsd_oracle as O
from tests.conftest import rel_err
from tests.test_driver import DRIVER, STEPS, read_tensor

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).parent / "golden"


def ccsd_step(cfg, extra=""):
    return f"""
- name: CoupledCluster
  in:
    method: Ccsd
    slicedEigenEnergies: EigenEnergies
    coulombIntegrals: CoulombIntegrals
    slicedCoulombVertex: CoulombVertex
    maxIterations: {cfg['max_iterations']}
    energyConvergence: {cfg['energy_convergence']}
    amplitudesConvergence: {cfg['amplitudes_convergence']}
    mixer:
      type: DiisMixer
      maxResidua: {cfg['max_residua']}
{extra}  out:
    energy: CcsdEnergy
    amplitudes: Amplitudes
"""


def run(workdir, text):
    (workdir / "cc4s.in").write_text(text)
    r = subprocess.run([str(DRIVER), "-i", "cc4s.in"], cwd=workdir,
                       capture_output=True, text=True, timeout=600)
    print(r.stdout[-3000:], flush=True)
    assert r.returncode == 0, r.stderr
    return yaml.safe_load((workdir / "cc4s.out.yaml").read_text())


def ccsd_out(report):
    for k in sorted(report["steps"]):
        if report["steps"][k]["name"] == "CoupledCluster":
            return report["steps"][k]["out"]
    raise AssertionError("no CoupledCluster step in the report")


def amplitudes(directory, name="Amplitudes"):
    h = yaml.safe_load((directory / f"{name}.yaml").read_text())
    assert h["type"] == "TensorSet"
    c = h["components"]
    return (read_tensor(directory, f"{name}.components.ph", c["ph"]),
            read_tensor(directory, f"{name}.components.pphh", c["pphh"]))


@pytest.mark.parametrize("golden,binary", [("real_no8_nv40_ng200", True),
                                           ("small_complex", False)],
                         ids=["real_no8_nv40_ng200_binary", "small_complex_text"])
def test_operation_file_matches_golden_and_oracle(tmp_path, golden, binary):
    gold = json.loads((GOLDEN / f"{golden}.json").read_text())
    c = gold["config"]
    print(f"driver {golden}", flush=True)
    inp = syn.generate(c["No"], c["Nv"], c["NG"], c["field"], c["seed"])
    syn.write_cc4s_inputs(inp, tmp_path, binary=binary)
    write = f"""
- name: Write
  in:
    source: Amplitudes
    fileName: "Amplitudes.yaml"
    binary: {1 if binary else 0}
  out: {{}}
"""
    rep = run(tmp_path, STEPS + ccsd_step(c) + write)
    out = ccsd_out(rep)
    assert out["convergenceReached"] == 1
    for key in ("correlation", "direct", "exchange", "secondOrder"):
        assert abs(out["energy"][key] - gold["energy"][key]) < 1e-9, key
    assert out["energy"]["unit"] == 1.0
    T1, T2 = amplitudes(tmp_path)
    P = O.CcsdProblem(inp.gamma, inp.eps, c["No"], c["field"])
    ref = O.coupled_cluster(P, max_iterations=c["max_iterations"],
                            energy_convergence=c["energy_convergence"],
                            amplitudes_convergence=c["amplitudes_convergence"],
                            max_residua=c["max_residua"])
    assert rel_err(T1, ref["amplitudes"][0]) < 1e-8
    assert rel_err(T2, ref["amplitudes"][1]) < 1e-8

    # restart from the written amplitudes (CoupledCluster.cxx:70-75)
    restart = STEPS + """
- name: Read
  in:
    fileName: "Amplitudes.yaml"
  out:
    destination: Start
""" + ccsd_step(c, "    initialAmplitudes: Start\n")
    rep2 = run(tmp_path, restart)
    out2 = ccsd_out(rep2)
    assert out2["convergenceReached"] == 1
    assert abs(out2["energy"]["correlation"] - gold["energy"]["correlation"]) < 1e-9
    assert rep2["steps"][6]["statistics"]["floatingPointOperations"] < \
        rep["steps"][5]["statistics"]["floatingPointOperations"]


def test_operation_file_mixed_precision(tmp_path):
    gold = json.loads((GOLDEN / "small_complex.json").read_text())
    c = gold["config"]
    inp = syn.generate(c["No"], c["Nv"], c["NG"], c["field"], c["seed"])
    syn.write_cc4s_inputs(inp, tmp_path)
    rep = run(tmp_path, STEPS + ccsd_step(c, "    precision: mixed\n    mixedSwitch: 1.0e-5\n"))
    out = ccsd_out(rep)
    assert out["convergenceReached"] == 1
    assert abs(out["energy"]["correlation"] - gold["energy"]["correlation"]) < 1e-8


def test_vertex_with_more_states_than_the_energies(tmp_path):
    """SliceOperator takes holes [0, No) and particles [Np - Nv, Np) of the
    vertex (SliceOperator.cxx:112-118): a vertex with extra states between
    them gives the energy of its hole/particle sub-vertex."""
    No, Nv, NG = 4, 12, 30
    inp = syn.generate(No, Nv, NG, "complex", 8)
    extra = 3
    Np = No + Nv
    big = np.zeros((NG, Np + extra, Np + extra), dtype=np.complex128, order="F")
    idx = list(range(No)) + list(range(No + extra, Np + extra))
    big[np.ix_(range(NG), idx, idx)] = inp.gamma
    wide = syn.CcsdInputs(gamma=big, eps=inp.eps, No=No, field="complex", scale=inp.scale)
    syn.write_cc4s_inputs(wide, tmp_path)
    # the energies tensor keeps No + Nv states
    cfg = dict(max_iterations=30, energy_convergence=1e-10, amplitudes_convergence=1e-8,
               max_residua=4)
    rep = run(tmp_path, STEPS + ccsd_step(cfg))
    P = O.CcsdProblem(inp.gamma, inp.eps, No, "complex")
    ref = O.coupled_cluster(P, max_iterations=30, energy_convergence=1e-10,
                            amplitudes_convergence=1e-8, max_residua=4)
    assert abs(ccsd_out(rep)["energy"]["correlation"] - ref["energy"]["correlation"]) < 1e-9


def test_reference_ueg_operation_file(tmp_path):
    """The reference's own operation file of test/tests/ueg/rs1.0-7occ-26virt
    (Drccd, Ccsd, PerturbativeTriplesReference) on the uniform-electron-gas
    inputs restated from UegVertexGenerator.cxx: every step reports the
    energies of its correct.out.yaml within 1e-9 Eh (the reference's check
    uses 1e-7)."""
    from oracle import ueg
    gamma, eps, fermi = ueg.generate(7, 26, 1.0)
    inp = syn.CcsdInputs(gamma=gamma, eps=eps, No=7, field="complex", scale=1.0,
                         fermi_energy=fermi)
    syn.write_cc4s_inputs(inp, tmp_path)
    s = ueg.UEG_RS1_NO7_NV26["settings"]
    common = f"""    slicedEigenEnergies: EigenEnergies
    coulombIntegrals: CoulombIntegrals
    slicedCoulombVertex: CoulombVertex
    maxIterations: {s['maxIterations']}
    energyConvergence: 1.0E-8
    amplitudesConvergence: 1.0E-8
    mixer:
      type: DiisMixer
      maxResidua: {s['maxResidua']}
"""
    ops = STEPS + f"""
- name: CoupledCluster
  in:
    method: Drccd
{common}  out:
    energy: CcsdEnergy

- name: CoupledCluster
  in:
    method: Ccsd
    integralsSliceSize: 100
{common}  out:
    energy: CcsdEnergy
    amplitudes: Amplitudes

- name: PerturbativeTriplesReference
  in:
    coulombIntegrals: CoulombIntegrals
    amplitudes: Amplitudes
    slicedEigenEnergies: EigenEnergies
  out:
    {{}}
"""
    rep = run(tmp_path, ops)
    for step, method in ((5, "Drccd"), (6, "Ccsd")):
        out = rep["steps"][step]["out"]
        assert out["convergenceReached"] == 1
        for key in ("correlation", "direct", "exchange", "secondOrder"):
            assert abs(out["energy"][key] - ueg.UEG_RS1_NO7_NV26[method][key]) < 1e-9, (method, key)
    triples = rep["steps"][7]
    assert triples["name"] == "PerturbativeTriplesReference"
    assert abs(triples["out"]["energy"]["correlation"]
               - ueg.UEG_RS1_NO7_NV26["Triples"]["correlation"]) < 1e-10


MP2_STEP = """
- name: SecondOrderPerturbationTheory
  in:
    coulombIntegrals: CoulombIntegrals
    slicedEigenEnergies: EigenEnergies
{fock}  out:
    energy: Mp2Energy
"""


def mp2_out(report):
    for k in sorted(report["steps"], key=int):
        if report["steps"][k]["name"] == "SecondOrderPerturbationTheory":
            return report["steps"][k]["out"]["energy"]


def test_second_order_ueg(tmp_path):
    """SecondOrderPerturbationTheory on the rs = 1, No = 7, Nv = 26 electron
    gas: the reference's recorded second-order energy (correct.out.yaml,
    CoupledCluster's secondOrder) and the oracle's direct and exchange terms."""
    from oracle import ueg
    gamma, eps, fermi = ueg.generate(7, 26, 1.0)
    print("second order, UEG rs=1 No=7 Nv=26", flush=True)
    inp = syn.CcsdInputs(gamma=gamma, eps=eps, No=7, field="complex", scale=1.0,
                         fermi_energy=fermi)
    syn.write_cc4s_inputs(inp, tmp_path)
    e = mp2_out(run(tmp_path, STEPS + MP2_STEP.format(fock="")))
    assert abs(e["secondOrder"] - ueg.UEG_RS1_NO7_NV26["Ccsd"]["secondOrder"]) < 1e-12
    want = O.second_order(O.CcsdProblem(gamma, eps, 7, "complex"))
    for key in ("direct", "exchange", "correlation"):
        assert abs(e[key] - want[key]) < 1e-12, key
    assert e["singles"] == 0.0


@pytest.mark.parametrize("field", ["real", "complex"])
def test_second_order_with_fock_singles(tmp_path, field):
    """A non-canonical Fock operator (ph couplings) read, sliced and passed as
    slicedFockOperator: every energy term matches the oracle."""
    from tests.test_driver import fock_operator
    No, Nv, NG = 5, 19, 40
    print(f"second order with Fock singles, {field}", flush=True)
    inp = syn.generate(No, Nv, NG, field, 13)
    syn.write_cc4s_inputs(inp, tmp_path)
    F, f = fock_operator(inp, 3)
    syn.write_tensor(tmp_path, "FockOperator", F, [(No + Nv, "State"), (No + Nv, "State")])
    fock = "    slicedFockOperator: FockOperator\n"
    steps = STEPS + """
- name: Read
  in:
    fileName: "FockOperator.yaml"
  out:
    destination: FockOperator
- name: SliceOperator
  in:
    slicedEigenEnergies: EigenEnergies
    operator: FockOperator
  out:
    slicedOperator: FockOperator
""" + MP2_STEP.format(fock=fock)
    e = mp2_out(run(tmp_path, steps))
    want = O.second_order(O.CcsdProblem(inp.gamma, inp.eps, No, field), fph=f)
    for key in ("singles", "direct", "exchange", "correlation"):
        assert abs(e[key] - want[key]) <= 1e-12 * abs(want["correlation"]), key
    assert e["secondOrder"] == e["correlation"]


def test_second_order_open_shell_spin_orbitals(tmp_path):
    """The spin-orbital restatement of a closed-shell system, marked open
    shell by State.yaml's Spin property (degeneracy 1), gives the closed-shell
    second-order energy."""
    from tests.test_oracle import _spin_orbitals
    No, Nv, NG = 3, 9, 20
    print("second order, open shell (spin orbitals)", flush=True)
    inp = syn.generate(No, Nv, NG, "complex", 17)
    g, e = _spin_orbitals(inp.gamma, np.asarray(inp.eps), No)
    so = syn.CcsdInputs(gamma=g, eps=e, No=2 * No, field="complex", scale=inp.scale,
                        fermi_energy=inp.fermi_energy)
    syn.write_cc4s_inputs(so, tmp_path)
    spins = [0] * No + [1] * No + [0] * Nv + [1] * Nv
    (tmp_path / "State.yaml").write_text(
        yaml.safe_dump({"propertyIndices": {"Spin": {i: s for i, s in enumerate(spins)}}}))
    got = mp2_out(run(tmp_path, STEPS + MP2_STEP.format(fock="")))
    want = O.second_order(O.CcsdProblem(inp.gamma, inp.eps, No, "complex"))
    assert abs(got["correlation"] - want["correlation"]) < 1e-12 * abs(want["correlation"])


@pytest.mark.parametrize("step", ["PerturbativeTriplesReference", "PerturbativeTriples"])
def test_triples_star_correction(tmp_path, step):
    """(T) with mp2PairEnergies on the rs = 1, No = 7, Nv = 26 electron gas:
    given 1.25 times the finite-basis MP2 pair energies, starCorrection is
    1.25 times the reference's recorded (T) energy (and the oracle's (T*))."""
    from oracle import ueg
    gamma, eps, fermi = ueg.generate(7, 26, 1.0)
    print(f"{step} with mp2PairEnergies, UEG", flush=True)
    inp = syn.CcsdInputs(gamma=gamma, eps=eps, No=7, field="complex", scale=1.0,
                         fermi_energy=fermi)
    syn.write_cc4s_inputs(inp, tmp_path)
    P = O.CcsdProblem(gamma, eps, 7, "complex")
    pairs = 1.25 * O.mp2_pair_energies(P).real
    syn.write_tensor(tmp_path, "Mp2PairEnergies", pairs, [(7, "State"), (7, "State")])
    s = ueg.UEG_RS1_NO7_NV26["settings"]
    ops = STEPS + f"""
- name: Read
  in:
    fileName: "Mp2PairEnergies.yaml"
  out:
    destination: Mp2PairEnergies
- name: CoupledCluster
  in:
    method: Ccsd
    slicedEigenEnergies: EigenEnergies
    coulombIntegrals: CoulombIntegrals
    slicedCoulombVertex: CoulombVertex
    integralsSliceSize: 100
    maxIterations: {s['maxIterations']}
    energyConvergence: 1.0E-8
    amplitudesConvergence: 1.0E-8
    mixer:
      type: DiisMixer
      maxResidua: {s['maxResidua']}
  out:
    amplitudes: Amplitudes
- name: {step}
  in:
    coulombIntegrals: CoulombIntegrals
    amplitudes: Amplitudes
    slicedEigenEnergies: EigenEnergies
    mp2PairEnergies: Mp2PairEnergies
  out:
    {{}}
"""
    rep = run(tmp_path, ops)
    e = rep["steps"][7]["out"]["energy"]
    ref_t = ueg.UEG_RS1_NO7_NV26["Triples"]["correlation"]
    assert abs(e["correlation"] - ref_t) < 1e-10
    assert abs(e["starCorrection"] - 1.25 * ref_t) < 1e-10
    assert abs(e["starCorrection"] - O.triples_star(P, e["correlation"], pairs)) < 1e-13
    keys = list(e)
    assert keys.index("starCorrection") < keys.index("unit") if step == "PerturbativeTriples" \
        else keys.index("starCorrection") > keys.index("unit")
