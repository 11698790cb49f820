"""The allocation-free energy entry point (cc4s_ccsd_energy_noalloc) against
cc4s_ccsd_energy, the context entry point it mirrors
(CoupledCluster::getEnergy, CoupledCluster.cxx:194-253): the same kernels
and reduction order, so the results agree to the last bit.  Two calls in a
row on one GPU share its scratch array; the second has a smaller input, and
a real-field call follows.  Device buffers are the caller's (allocated here,
in the test, through the HIP runtime the library uses)."""

import ctypes

import numpy as np
import pytest

from cc4s_amd import api
from cc4s_amd import synthetic as syn

pytestmark = pytest.mark.gpu


class _DeviceArray:
    def __init__(self, a):
        api.load_library()
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.ptr = ctypes.c_void_p()
        self.nbytes = a.nbytes
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(a.nbytes)) == 0
        assert self.hip.hipMemcpy(self.ptr, ctypes.c_void_p(a.ctypes.data),
                                  ctypes.c_size_t(a.nbytes), 1) == 0

    def free(self):
        assert self.hip.hipFree(self.ptr) == 0


def _case(No, Nv, NG, field, seed):
    """(context energy, allocation-free energy) of seeded amplitudes."""
    inp = syn.generate(No, Nv, NG, field, seed=seed)
    dtype = np.complex128 if field == "complex" else np.float64
    T1, T2 = syn.random_amplitudes(No, Nv, field, seed=seed + 1)
    with api.CoupledCluster(No, Nv, NG, field=field) as cc:
        cc.set_inputs(inp.gamma, inp.eps)
        _, V = cc.residuum(None)            # first-pass residuum: R2 = V_abij
        ref = cc.energy((T1, T2))
    T = np.concatenate([np.asarray(T1, dtype).ravel(order="F"),
                        np.asarray(T2, dtype).ravel(order="F")])
    V = np.asfortranarray(np.asarray(V, dtype))
    dT, dV = _DeviceArray(T), _DeviceArray(V)
    try:
        got = api.energy_noalloc(field, No, Nv, dT.ptr.value, dV.ptr.value)
    finally:
        dT.free()
        dV.free()
    return ref, got


@pytest.mark.parametrize("field", ["complex", "real"])
def test_noalloc_energy_equals_the_context_energy(field):
    for No, Nv, NG, seed in ((8, 40, 120, 5), (4, 16, 40, 9)):   # second call smaller
        print(f"noalloc energy: field {field} No={No} Nv={Nv} NG={NG} seed {seed}",
              flush=True)
        ref, got = _case(No, Nv, NG, field, seed)
        print(f"  context {ref[0]!r}  noalloc {got[0]!r}", flush=True)
        for r, g in zip(ref, got):
            assert g == r, (ref, got)        # same kernels, same reduction order


def test_noalloc_energy_rejects_bad_arguments():
    with pytest.raises(api.CcsdError, match="null argument"):
        api.energy_noalloc("complex", 4, 16, 0, 0)
    with pytest.raises(api.CcsdError, match="positive"):
        api.energy_noalloc("complex", 0, 16, 8, 8)
    with pytest.raises(api.CcsdError, match="too large"):
        api.energy_noalloc("complex", 1 << 21, 16, 8, 8)
