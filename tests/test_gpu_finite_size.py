"""FiniteSizeCorrection on the GPU: the transition structure factor through the
C ABI against the oracle, and the driver's FiniteSizeCorrection step end to
end on the uniform-electron-gas inputs restated from UegVertexGenerator.cxx.
"""

import pathlib
import subprocess

import numpy as np
import pytest
import yaml

from cc4s_amd import api
from cc4s_amd import synthetic as syn
from oracle import finite_size as F
from oracle import ueg
from tests.conftest import rel_err
from tests.test_driver import DRIVER, STEPS, read_tensor

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("with_u", [False, True], ids=["on_grid", "singular_vectors"])
def test_structure_factor_parity(with_u):
    No, Nv, NG = 4, 12, 20
    print(f"transition structure factor, singular vectors {with_u}", flush=True)
    inp = syn.generate(No, Nv, NG, "complex", 9)
    T1, T2 = syn.random_amplitudes(No, Nv, "complex", 4, t1_scale=0.05, t2_scale=0.05)
    rng = np.random.default_rng(6)
    NGg = 27 if with_u else NG
    U = (rng.standard_normal((NGg, NG)) + 1j * rng.standard_normal((NGg, NG))) if with_u else None
    V = rng.uniform(0.5, 2.0, NGg)
    with api.CoupledCluster(No, Nv, NG) as cc:
        cc.set_inputs(inp.gamma, inp.eps)
        cc.set_amplitudes(T1, T2)
        S = cc.transition_structure_factor(V, U)
    want = F.transition_structure_factor(inp.gamma, No, T1, T2, V, U)
    assert rel_err(S, want) < 1e-12


def test_structure_factor_refuses_the_real_field():
    No, Nv, NG = 2, 5, 6
    inp = syn.generate(No, Nv, NG, "real", 1)
    with api.CoupledCluster(No, Nv, NG, field="real") as cc:
        cc.set_inputs(inp.gamma, inp.eps)
        cc.set_amplitudes(np.zeros((Nv, No)), np.zeros((Nv, Nv, No, No)))
        with pytest.raises(api.CcsdError, match="complex field"):
            cc.transition_structure_factor(np.ones(NG))


def test_driver_finite_size_correction_ueg(tmp_path):
    """CCSD on the rs = 1, No = 7, Nv = 26 electron gas, then the
    FiniteSizeCorrection step (no singular vectors: the vertex is on the
    momentum grid); the report's energies and the stored structure factor
    match the oracle on the written amplitudes."""
    No, Nv, N = 7, 26, 6
    gamma, eps, fermi = ueg.generate(No, Nv, 1.0)
    grid, B, V = ueg.momentum_grid(No, Nv, 1.0)
    inp = syn.CcsdInputs(gamma=gamma, eps=eps, No=No, field="complex", scale=1.0,
                         fermi_energy=fermi)
    syn.write_cc4s_inputs(inp, tmp_path)
    NG = grid.shape[0]
    syn.write_tensor(tmp_path, "GridVectors", grid.T, [(3, "Vector"), (NG, "Momentum")],
                     meta={"Gi": B[0].tolist(), "Gj": B[1].tolist(), "Gk": B[2].tolist()})
    syn.write_tensor(tmp_path, "CoulombPotential", V, [(NG, "Momentum")])
    ops = STEPS + f"""
- name: Read
  in:
    fileName: "GridVectors.yaml"
  out:
    destination: GridVectors
- name: Read
  in:
    fileName: "CoulombPotential.yaml"
  out:
    destination: CoulombPotential
- name: CoupledCluster
  in:
    method: Ccsd
    slicedEigenEnergies: EigenEnergies
    coulombIntegrals: CoulombIntegrals
    slicedCoulombVertex: CoulombVertex
    maxIterations: 30
    energyConvergence: 1.0E-10
    amplitudesConvergence: 1.0E-9
  out:
    energy: CcsdEnergy
    amplitudes: Amplitudes
- name: FiniteSizeCorrection
  in:
    slicedCoulombVertex: CoulombVertex
    amplitudes: Amplitudes
    coulombPotential: CoulombPotential
    gridVectors: GridVectors
    interpolationGridSize: {N}
  out:
    transitionStructureFactor: SF
- name: Write
  in:
    source: Amplitudes
    fileName: "Amplitudes.yaml"
  out: {{}}
- name: Write
  in:
    source: SF
    fileName: "SF.yaml"
  out: {{}}
"""
    (tmp_path / "cc4s.in").write_text(ops)
    r = subprocess.run([str(DRIVER), "-i", "cc4s.in"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=600)
    print(r.stdout[-2000:], flush=True)
    assert r.returncode == 0, r.stderr
    rep = yaml.safe_load((tmp_path / "cc4s.out.yaml").read_text())
    steps = [rep["steps"][k] for k in sorted(rep["steps"], key=int)]
    fsc = next(s for s in steps if s["name"] == "FiniteSizeCorrection")
    h = yaml.safe_load((tmp_path / "Amplitudes.yaml").read_text())["components"]
    T1 = read_tensor(tmp_path, "Amplitudes.components.ph", h["ph"])
    T2 = read_tensor(tmp_path, "Amplitudes.components.pphh", h["pphh"])
    S_ref = F.transition_structure_factor(gamma, No, T1, T2, V)
    S = read_tensor(tmp_path, "SF", yaml.safe_load((tmp_path / "SF.yaml").read_text()))
    assert rel_err(S, S_ref) < 1e-12
    corrected, uncorrected = F.finite_size_correction(grid, B, S_ref, N=N)
    e = fsc["out"]["energy"]
    assert abs(e["uncorrected"] - uncorrected) <= 1e-10 * abs(uncorrected)
    assert abs(e["corrected"] - corrected) <= 1e-10 * abs(corrected)
    assert abs(e["correction"] - (corrected - uncorrected)) <= 1e-10 * abs(corrected)
    # the structure factor decomposes the correlation energy over G
    ccsd = next(s for s in steps if s["name"] == "CoupledCluster")
    assert abs(np.sum(V * S) - ccsd["out"]["energy"]["correlation"]) < 1e-10


def test_driver_finite_size_correction_real_amplitudes(tmp_path):
    """Real Coulomb integrals (halfGrid) give real amplitudes; the
    FiniteSizeCorrection step takes them to the complex field
    (FiniteSizeCorrection.cxx:103-117) and brings the vertex to a momentum
    grid with singular vectors [G, F]; structure factor and energies match
    the oracle."""
    from tests.test_finite_size import _grid, _lattice
    No, Nv, NF, N = 4, 12, 20, 3
    print("finite-size correction, real amplitudes with singular vectors", flush=True)
    inp = syn.generate(No, Nv, NF, "real", 31)
    syn.write_cc4s_inputs(inp, tmp_path)
    B = _lattice("bcc")
    grid = _grid(B, 2)
    NGg = grid.shape[0]
    g2 = np.sum(grid * grid, axis=1)
    V = np.where(g2 > 0, 4.0 * np.pi / np.where(g2 > 0, g2, 1.0), 2.5)
    rng = np.random.default_rng(12)
    U = (rng.standard_normal((NGg, NF)) + 1j * rng.standard_normal((NGg, NF))) / np.sqrt(NF)
    syn.write_tensor(tmp_path, "GridVectors", grid.T, [(3, "Vector"), (NGg, "Momentum")],
                     meta={"Gi": B[0].tolist(), "Gj": B[1].tolist(), "Gk": B[2].tolist()})
    syn.write_tensor(tmp_path, "CoulombPotential", V, [(NGg, "Momentum")])
    syn.write_tensor(tmp_path, "SingularVectors", U, [(NGg, "Momentum"), (NF, "AuxiliaryField")])
    ops = STEPS + f"""
- name: Read
  in:
    fileName: "GridVectors.yaml"
  out:
    destination: GridVectors
- name: Read
  in:
    fileName: "CoulombPotential.yaml"
  out:
    destination: CoulombPotential
- name: Read
  in:
    fileName: "SingularVectors.yaml"
  out:
    destination: SingularVectors
- name: CoupledCluster
  in:
    method: Ccsd
    slicedEigenEnergies: EigenEnergies
    coulombIntegrals: CoulombIntegrals
    slicedCoulombVertex: CoulombVertex
    maxIterations: 4
  out:
    amplitudes: Amplitudes
- name: FiniteSizeCorrection
  in:
    slicedCoulombVertex: CoulombVertex
    amplitudes: Amplitudes
    coulombVertexSingularVectors: SingularVectors
    coulombPotential: CoulombPotential
    gridVectors: GridVectors
    interpolationGridSize: {N}
  out:
    transitionStructureFactor: SF
- name: Write
  in:
    source: Amplitudes
    fileName: "Amplitudes.yaml"
  out: {{}}
- name: Write
  in:
    source: SF
    fileName: "SF.yaml"
  out: {{}}
"""
    (tmp_path / "cc4s.in").write_text(ops)
    r = subprocess.run([str(DRIVER), "-i", "cc4s.in"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=600)
    print(r.stdout[-2000:], flush=True)
    assert r.returncode == 0, r.stderr
    rep = yaml.safe_load((tmp_path / "cc4s.out.yaml").read_text())
    steps = [rep["steps"][k] for k in sorted(rep["steps"], key=int)]
    fsc = next(s for s in steps if s["name"] == "FiniteSizeCorrection")
    h = yaml.safe_load((tmp_path / "Amplitudes.yaml").read_text())["components"]
    assert h["pphh"]["scalarType"] == "Real64"
    T1 = read_tensor(tmp_path, "Amplitudes.components.ph", h["ph"]).astype(complex)
    T2 = read_tensor(tmp_path, "Amplitudes.components.pphh", h["pphh"]).astype(complex)
    S_ref = F.transition_structure_factor(inp.gamma, No, T1, T2, V, U)
    S = read_tensor(tmp_path, "SF", yaml.safe_load((tmp_path / "SF.yaml").read_text()))
    assert rel_err(S, S_ref) < 1e-12
    corrected, uncorrected = F.finite_size_correction(grid, B, S_ref, N=N)
    e = fsc["out"]["energy"]
    assert abs(e["uncorrected"] - uncorrected) <= 1e-10 * abs(uncorrected)
    assert abs(e["corrected"] - corrected) <= 1e-10 * abs(corrected)
