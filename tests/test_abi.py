"""The C-ABI library loads and exports every entry point of
include/cc4s_ccsd.h (no GPU needed: nothing is computed)."""

import ctypes
import pathlib
import re

import pytest

from cc4s_amd import api

ROOT = pathlib.Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "cc4s_ccsd.h").read_text()
    return sorted(set(re.findall(r"\b(cc4s_\w+)\s*\(", text)))


def test_library_built_and_loadable():
    assert api.LIB_PATH.exists(), "run python -m cc4s_amd.build"
    api.load_library()


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(str(api.LIB_PATH))
    declared = _declared()
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(lib, name), f"{name} missing from the library"
    assert sorted(api.EXPORTS) == declared


def test_default_params_are_the_reference_defaults():
    p = api.default_params()
    # CoupledCluster.hpp:45-49, DiisMixer.cxx:33, Ccsd.cxx:95
    assert p.max_iterations == 16
    assert p.energy_convergence == 1e-7
    assert p.amplitudes_convergence == 1e-6
    assert p.level_shift == 0.0
    assert p.max_residua == 4
    assert p.ppl == 1
    assert p.integrals_slice_size == 0
    assert p.field == api.FIELD_COMPLEX
    # mixed-precision settings (no counterpart in the reference)
    assert p.precision == api.PRECISION_FP64
    assert p.mixed_switch == 1e-5
    assert p.mixed_refine == 1 and p.mixed_split == 1 and p.mixed_max_k == 2048


def test_unsupported_precision_is_refused_before_touching_a_device():
    with pytest.raises(api.CcsdError, match="precision mode"):
        api.CoupledCluster(2, 4, 3, precision=7)
    with pytest.raises(api.CcsdError, match="complex field"):
        api.CoupledCluster(2, 4, 3, field="real", precision=api.PRECISION_MIXED)


def test_single_occupied_orbital_is_refused_before_touching_a_device():
    with pytest.raises(api.CcsdError, match="No = 1"):
        api.CoupledCluster(1, 9, 6, field="complex")
    with pytest.raises(api.CcsdError, match="unknown ladder"):
        api.CoupledCluster(2, 9, 6, field="complex", ladder="tiled")



@pytest.mark.parametrize("cname,pyclass", [("cc4s_ccsd_params", "Params"),
                                           ("cc4s_ccsd_iteration", "Iteration"),
                                           ("cc4s_ccsd_result", "Result")])
def test_ctypes_structs_follow_the_header(cname, pyclass):
    """The ctypes mirrors list the header's struct members in its order."""
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "cc4s_ccsd.h").read_text(),
                  flags=re.S)
    body = re.search(r"typedef struct \{([^{}]*)\}\s*" + cname + ";", text).group(1)
    names = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        parts = decl.split(None, 1)
        names += [n.strip().lstrip("*") for n in parts[1].split(",")]
    assert [f for f, _ in getattr(api, pyclass)._fields_] == names


def test_amplitude_and_energy_shapes_are_checked_before_the_library_reads_them():
    """The library reads Nv*No, Nv^2*No^2 and No+Nv values; a short or
    transposed array is refused in Python (no device call is made)."""
    import numpy as np
    cc = object.__new__(api.CoupledCluster)
    cc.No, cc.Nv, cc.NG, cc.dtype = 2, 3, 4, np.complex128
    cc._amplitude_arrays(np.zeros((3, 2)), np.zeros((3, 3, 2, 2)))
    with pytest.raises(ValueError, match="T1 shape"):
        cc._amplitude_arrays(np.zeros((2, 3)), np.zeros((3, 3, 2, 2)))
    with pytest.raises(ValueError, match="T2 shape"):
        cc._amplitude_arrays(np.zeros((3, 2)), np.zeros((3, 3, 2)))
    with pytest.raises(ValueError, match="eps shape"):
        cc._eps(np.zeros(4))
    assert cc._eps(np.zeros(5)).shape == (5,)
