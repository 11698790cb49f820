"""The cc4s operation-file driver (cc4s_amd/cc4s_ccsd) on the CPU.

Runs the driver in dry-run mode on cc4s-format files written by the synthetic
generator: file parsing (YAML headers, binary and text elements), the steps
Read / DefineHolesAndParticles / SliceOperator / VertexCoulombIntegrals /
Write, the field choice by halfGrid, the report, and the reference's error
texts.  The CoupledCluster step itself needs the GPU (tests/test_gpu_driver.py).
"""

import pathlib
import subprocess

import numpy as np
import pytest
import yaml

from cc4s_amd import build as B
from cc4s_amd import synthetic as syn
from oracle import ccsd_oracle as O

DRIVER = B.DRIVER

STEPS = """
- name: Read
  in:
    fileName: "EigenEnergies.yaml"
  out:
    destination: EigenEnergies
- name: Read
  in:
    fileName: "CoulombVertex.yaml"   # the vertex
  out:
    destination: CoulombVertex
- name: DefineHolesAndParticles
  in:
    eigenEnergies: EigenEnergies
  out:
    slicedEigenEnergies: EigenEnergies
- name: SliceOperator
  in:
    slicedEigenEnergies: EigenEnergies
    operator: CoulombVertex
  out:
    slicedOperator: CoulombVertex
- name: VertexCoulombIntegrals
  in:
    slicedCoulombVertex: CoulombVertex
  out:
    coulombIntegrals: CoulombIntegrals
"""

CCSD = """
- name: CoupledCluster
  in:
    method: Ccsd
    slicedEigenEnergies: EigenEnergies
    coulombIntegrals: CoulombIntegrals
    slicedCoulombVertex: CoulombVertex
    maxIterations: 20
    energyConvergence: 1.0E-8
    mixer: {type: DiisMixer, maxResidua: 4}
  out:
    energy: CcsdEnergy
    amplitudes: Amplitudes
"""


def run_driver(workdir, text, *extra):
    (workdir / "cc4s.in").write_text(text)
    return subprocess.run([str(DRIVER), "-i", "cc4s.in", *extra], cwd=workdir,
                          capture_output=True, text=True, timeout=120)


def read_tensor(directory, base, header):
    """Elements of a tensor written by the driver (binary or text)."""
    lens = [d["length"] for _, d in sorted(header["dimensions"].items())]
    cplx = header["scalarType"] == "Complex64"
    path = pathlib.Path(directory) / f"{base}.elements"
    if header["elements"]["type"] == "IeeeBinaryFile":
        v = np.fromfile(path, dtype=np.complex128 if cplx else np.float64)
    else:
        rows = path.read_text().split()
        if cplx:
            v = np.array([complex(*map(float, r.strip("()").split(","))) for r in rows])
        else:
            v = np.array([float(r) for r in rows])
    return v.reshape(lens, order="F")


@pytest.fixture(scope="module")
def built():
    if not DRIVER.exists():
        pytest.skip("driver not built (python -m cc4s_amd.build)")
    return DRIVER


@pytest.mark.parametrize("binary", [True, False], ids=["binary", "text"])
@pytest.mark.parametrize("field", ["real", "complex"])
def test_dry_run_reads_and_slices(built, tmp_path, field, binary):
    inp = syn.generate(3, 7, 5, field, 4)
    syn.write_cc4s_inputs(inp, tmp_path, binary=binary)
    r = run_driver(tmp_path, STEPS + CCSD + """
- name: Write
  in:
    source: EigenEnergies
    fileName: "eps.yaml"
    binary: 1
  out: {}
""", "--dry-run")
    assert r.returncode == 0, r.stderr
    assert "number of holes     No: 3" in r.stdout
    assert ("Using real Coulomb integrals" if field == "real"
            else "Using complex Coulomb integrals") in r.stdout
    rep = yaml.safe_load((tmp_path / "dry-cc4s.out.yaml").read_text())
    names = [rep["steps"][i]["name"] for i in range(7)]
    assert names == ["Read", "Read", "DefineHolesAndParticles", "SliceOperator",
                     "VertexCoulombIntegrals", "CoupledCluster", "Write"]
    cc_in = rep["steps"][5]["in"]
    # defaults reported as the reference's argument map does
    assert cc_in["levelShift"] == 0.0 and cc_in["ppl"] == 1
    assert cc_in["integralsSliceSize"] == 3
    # the written sliced eigen energies (a TensorSet {h, p}) hold eps
    out = yaml.safe_load((tmp_path / "eps.yaml").read_text())
    assert out["type"] == "TensorSet"
    h = read_tensor(tmp_path, "eps.components.h", out["components"]["h"])
    p = read_tensor(tmp_path, "eps.components.p", out["components"]["p"])
    assert np.array_equal(np.concatenate([h, p]), inp.eps)


def test_tensor_round_trip_text_and_binary(built, tmp_path):
    """Read then Write of the vertex, in text and binary, reproduces the bytes
    of the values (17 significant digits in text)."""
    inp = syn.generate(2, 5, 3, "complex", 9)
    syn.write_cc4s_inputs(inp, tmp_path, binary=False)
    (tmp_path / "sub").mkdir()
    r = run_driver(tmp_path, """
- name: Read
  in: {fileName: "CoulombVertex.yaml"}
  out: {destination: V}
- name: Write
  in: {source: V, fileName: "sub/Vb.yaml", binary: 1}
  out: {}
- name: Write
  in: {source: V, fileName: "Vt.yaml"}
  out: {}
""", "--dry-run")
    assert r.returncode == 0, r.stderr
    hb = yaml.safe_load((tmp_path / "sub" / "Vb.yaml").read_text())
    ht = yaml.safe_load((tmp_path / "Vt.yaml").read_text())
    assert hb["scalarType"] == "Complex64" and hb["metaData"]["halfGrid"] == 0
    vb = read_tensor(tmp_path / "sub", "Vb", hb)
    vt = read_tensor(tmp_path, "Vt", ht)
    assert np.array_equal(vb, inp.gamma) and np.array_equal(vt, inp.gamma)


def test_flow_value_on_the_next_line(built, tmp_path):
    """`out:` followed by an indented `{}` (the layout of the reference's own
    operation files, e.g. test/tests/ueg/rs1.0-7occ-26virt/cc4s.in:69-70)."""
    inp = syn.generate(2, 5, 3, "complex", 9)
    syn.write_cc4s_inputs(inp, tmp_path)
    r = run_driver(tmp_path, """
- name: Read
  in:
    fileName: "CoulombVertex.yaml"
  out:
    destination: V
- name: Write
  in:
    source: V
    fileName: "Vt.yaml"
  out:
    {}
""", "--dry-run")
    assert r.returncode == 0, r.stderr
    assert yaml.safe_load((tmp_path / "Vt.yaml").read_text())["scalarType"] == "Complex64"


@pytest.mark.parametrize("case,message", [
    ("unknown", "unknown algorithm: BasisSetCorrection"),
    ("fermi_high", "Fermi energy above all eigen energies."),
    ("fermi_low", "Fermi energy below all eigen energies."),
    ("version", "serialization versions mismatch"),
    ("method", "Unknown method: 'Ccsdt'"),
    ("mixer", "Unknown mixer type: 'LinearMixer'"),
    ("ladder", "unknown ladder 'tiled' (slab | fused | paired)"),
])
def test_reference_errors(built, tmp_path, case, message):
    inp = syn.generate(2, 5, 3, "complex", 1)
    syn.write_cc4s_inputs(inp, tmp_path)
    steps = STEPS + CCSD
    if case == "unknown":
        steps += "\n- name: BasisSetCorrection\n  in: {}\n  out: {}\n"
    if case in ("fermi_high", "fermi_low"):
        h = yaml.safe_load((tmp_path / "EigenEnergies.yaml").read_text())
        h["metaData"]["fermiEnergy"] = 10.0 if case == "fermi_high" else -10.0
        (tmp_path / "EigenEnergies.yaml").write_text(yaml.safe_dump(h))
    if case == "version":
        h = yaml.safe_load((tmp_path / "CoulombVertex.yaml").read_text())
        h["version"] = 99
        (tmp_path / "CoulombVertex.yaml").write_text(yaml.safe_dump(h))
    if case == "method":
        steps = steps.replace("method: Ccsd", "method: Ccsdt")
    if case == "mixer":
        steps = steps.replace("type: DiisMixer", "type: LinearMixer")
    if case == "ladder":
        steps = steps.replace("    maxIterations: 20\n", "    maxIterations: 20\n    ladder: tiled\n")
    r = run_driver(tmp_path, steps, "--dry-run")
    assert r.returncode == 1
    assert message in r.stderr


MP2 = """
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
- name: SecondOrderPerturbationTheory
  in:
    coulombIntegrals: CoulombIntegrals
    slicedEigenEnergies: EigenEnergies
    slicedFockOperator: FockOperator
  out:
    energy: Mp2Energy
- name: Write
  in:
    source: FockOperator
    fileName: "SlicedFock.yaml"
  out: {}
"""


def fock_operator(inp, seed, scale=0.05):
    """A Hermitian Fock operator [State, State]: the eigen energies on the
    diagonal, random hole-particle couplings (cc4s layout F[p, q])."""
    No, Np = inp.No, inp.Np
    rng = np.random.default_rng(seed)
    cplx = inp.field == "complex"
    F = np.diag(np.asarray(inp.eps, dtype=complex if cplx else float))
    f = scale * rng.standard_normal((Np - No, No))
    if cplx:
        f = f + 1j * scale * rng.standard_normal((Np - No, No))
    F[No:, :No] = f
    F[:No, No:] = np.conj(f).T
    return F, f


@pytest.mark.parametrize("spin", ["closed", "open"])
@pytest.mark.parametrize("field", ["real", "complex"])
def test_second_order_singles_and_fock_slices(built, tmp_path, field, spin):
    """SliceOperator of a [State, State] operator gives {hh, hp, ph, pp};
    SecondOrderPerturbationTheory's singles term (host, also in a dry run)
    equals the oracle's, halved when State.yaml's Spin property marks an
    open shell (SecondOrderPerturbationTheory.cxx:67-85)."""
    No, Nv, NG = 3, 7, 5
    inp = syn.generate(No, Nv, NG, field, 6)
    syn.write_cc4s_inputs(inp, tmp_path)
    F, f = fock_operator(inp, 2)
    syn.write_tensor(tmp_path, "FockOperator", F, [(No + Nv, "State"), (No + Nv, "State")])
    if spin == "open":
        spins = {i: (0 if i < (No + Nv) // 2 else 1) for i in range(No + Nv)}
        (tmp_path / "State.yaml").write_text(yaml.safe_dump({"propertyIndices": {"Spin": spins}}))
    r = run_driver(tmp_path, STEPS + MP2, "--dry-run")
    assert r.returncode == 0, r.stderr
    rep = yaml.safe_load((tmp_path / "dry-cc4s.out.yaml").read_text())
    e = rep["steps"][7]["out"]["energy"]
    P = O.CcsdProblem(inp.gamma, inp.eps, No, field)
    want = O.second_order(P, fph=f, degeneracy=1.0 if spin == "open" else 2.0)["singles"]
    assert abs(e["singles"] - want) < 1e-14 * abs(want)
    assert e["secondOrderSingles"] == e["singles"]
    out = yaml.safe_load((tmp_path / "SlicedFock.yaml").read_text())
    rng = {"h": slice(0, No), "p": slice(No, No + Nv)}
    for part in ("hh", "hp", "ph", "pp"):
        got = read_tensor(tmp_path, f"SlicedFock.components.{part}", out["components"][part])
        assert np.array_equal(got, F[rng[part[0]], rng[part[1]]]), part


@pytest.mark.parametrize("case,message", [
    ("fock_not_sliced", "expecting a TensorSet {hh, hp, ph, pp} as 'slicedFockOperator'"),
    ("fock_missing", "'slicedFockOperator' does not name a stored object"),
    ("pairs_missing", "'mp2PairEnergies' does not name a stored object"),
    ("cT", "cT (complete triples) is not supported"),
])
def test_second_order_and_triples_argument_errors(built, tmp_path, case, message):
    """Argument errors of SecondOrderPerturbationTheory and
    PerturbativeTriples stop the run before any GPU work (dry run)."""
    inp = syn.generate(2, 5, 3, "complex", 1)
    syn.write_cc4s_inputs(inp, tmp_path)
    if case.startswith("fock"):
        fock = "EigenEnergies" if case == "fock_not_sliced" else "NoSuchFock"
        steps = STEPS + f"""
- name: SecondOrderPerturbationTheory
  in:
    coulombIntegrals: CoulombIntegrals
    slicedEigenEnergies: EigenEnergies
    slicedFockOperator: {fock}
  out: {{}}
"""
    else:
        extra = "    mp2PairEnergies: NoSuchPairs\n" if case == "pairs_missing" else "    cT: 1\n"
        # any stored object passes the name lookup; the argument checks
        # come before the amplitudes are inspected
        steps = STEPS + f"""
- name: PerturbativeTriples
  in:
    coulombIntegrals: CoulombIntegrals
    amplitudes: EigenEnergies
    slicedEigenEnergies: EigenEnergies
{extra}  out: {{}}
"""
    r = run_driver(tmp_path, steps, "--dry-run")
    assert r.returncode == 1
    assert message in r.stderr
