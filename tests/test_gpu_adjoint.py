"""Energy adjoint on the GPU through the C ABI against the oracle.

Bar: FP64 forward energy and gradient within 1e-12 relative of the oracle's
``vertex_energy`` / ``vertex_gradient``; FP32 gradient and structure factor
within 1e-5 relative (FP32 operands and accumulation); central finite
differences of the GPU forward (step 1e-5 in Gamma) agree with the GPU
gradient; sum_G S(G) equals the energy to 1e-12 relative.
"""

import numpy as np
import pytest

from cc4s_amd import api
from cc4s_amd import synthetic as syn
from oracle import ccsd_oracle as O
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def _amps(No, Nv, field, seed, scale=0.05):
    T1, T2 = syn.random_amplitudes(No, Nv, field, seed, t1_scale=scale,
                                   t2_scale=scale)
    if field == "real":
        return T1.real.copy(), T2.real.copy()
    return T1, T2


def _vertex(No, Nv, NG, field, seed, hermitian):
    g = syn.generate(No, Nv, NG, field, seed).gamma
    if not hermitian:
        rng = np.random.default_rng(seed)
        g = g + 0.003 * (rng.standard_normal(g.shape) + 1j * rng.standard_normal(g.shape))
    return np.asfortranarray(g)


@pytest.mark.parametrize("field", ["complex", "real"])
@pytest.mark.parametrize("hermitian", [True, False], ids=["herm", "nonherm"])
@pytest.mark.parametrize("shape", [(3, 7, 5), (4, 12, 30), (8, 40, 200)],
                         ids=["o3v7G5", "o4v12G30", "o8v40G200"])
def test_forward_backward_parity(field, hermitian, shape):
    No, Nv, NG = shape
    print(f"adjoint parity {field} hermitian={hermitian} {shape}", flush=True)
    g = _vertex(No, Nv, NG, field, 31, hermitian)
    eps = syn.eigen_energies(No, Nv, 31)
    T = _amps(No, Nv, field, 17)
    with api.CoupledCluster(No, Nv, NG, field=field) as cc:
        cc.set_inputs(g, eps)
        cc.set_amplitudes(*T)
        E_ref = O.vertex_energy(g, T, No, field)
        g_ref, S_ref = O.vertex_gradient(g, T, No, field)
        E = cc.energy_forward()
        assert abs(E - E_ref) <= 1e-12 * abs(E_ref)
        assert abs(E - cc.energy(T)[0].real) <= 1e-12 * abs(E_ref)
        grad, S = cc.energy_backward()
        assert rel_err(grad, g_ref) < 1e-12
        assert rel_err(S, S_ref) < 1e-12
        assert abs(np.sum(S) - E) <= 1e-12 * abs(E)
        grad32, S32 = cc.energy_backward(precision=api.PRECISION_FP32)
        assert rel_err(grad32, g_ref) < 1e-5
        assert rel_err(S32, S_ref) < 1e-5
        # a vertex passed in host memory (not the input vertex)
        g2 = np.asfortranarray(1.01 * g)
        assert abs(cc.energy_forward(g2) - O.vertex_energy(g2, T, No, field)) \
            <= 1e-12 * abs(E_ref)
        grad2, _ = cc.energy_backward(g2, grad=True)
        assert rel_err(grad2, 1.01 * g_ref) < 1e-12


@pytest.mark.parametrize("field", ["complex", "real"])
def test_finite_differences_of_forward(field):
    No, Nv, NG = 4, 12, 30
    g = _vertex(No, Nv, NG, field, 5, True)
    eps = syn.eigen_energies(No, Nv, 5)
    rng = np.random.default_rng(6)
    with api.CoupledCluster(No, Nv, NG, field=field, maxIterations=4) as cc:
        cc.set_inputs(g, eps)
        cc.run()
        grad, S = cc.energy_backward()
        h = 1e-5
        Np = No + Nv
        # entries of the hp and ph blocks, both real components
        for _ in range(12):
            G = int(rng.integers(NG))
            if rng.random() < 0.5:
                p, q = int(rng.integers(No)), No + int(rng.integers(Nv))
            else:
                p, q = No + int(rng.integers(Nv)), int(rng.integers(No))
            for unit, part in ((1.0, grad[G, p, q].real), (1j, grad[G, p, q].imag)):
                gp = g.copy()
                gm = g.copy()
                gp[G, p, q] += h * unit
                gm[G, p, q] -= h * unit
                fd = (cc.energy_forward(gp) - cc.energy_forward(gm)) / (2 * h)
                assert abs(fd - part) < 1e-7 * np.max(np.abs(grad)) + 1e-9
        # G slices: E(slice scaled by 1 + t) = E_rest + (1 + t)^2 S(G)
        for G in rng.choice(NG, size=5, replace=False):
            gp = g.copy()
            gm = g.copy()
            gp[G] *= 1 + h
            gm[G] *= 1 - h
            fd = (cc.energy_forward(gp) - cc.energy_forward(gm)) / (4 * h)
            assert abs(fd - S[G]) < 1e-7 * np.max(np.abs(S))
        assert np.all(grad[:, :No, :No] == 0) and np.all(grad[:, No:, No:] == 0)
        assert Np == g.shape[1]


class _DeviceArray:
    """Device memory through the HIP runtime the library itself uses."""

    def __init__(self, nbytes):
        import ctypes
        api.load_library()
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.ptr = ctypes.c_void_p()
        self.nbytes = nbytes
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(nbytes)) == 0

    def upload(self, a):
        import ctypes
        assert a.nbytes == self.nbytes
        assert self.hip.hipMemcpy(self.ptr, ctypes.c_void_p(a.ctypes.data),
                                  ctypes.c_size_t(self.nbytes), 1) == 0

    def download(self, a):
        import ctypes
        assert a.nbytes == self.nbytes
        assert self.hip.hipMemcpy(ctypes.c_void_p(a.ctypes.data), self.ptr,
                                  ctypes.c_size_t(self.nbytes), 2) == 0

    def free(self):
        self.hip.hipFree(self.ptr)


def test_device_buffers():
    """Vertex and gradient in device memory give the host-buffer results."""
    No, Nv, NG = 4, 12, 30
    g = _vertex(No, Nv, NG, "complex", 9, True)
    eps = syn.eigen_energies(No, Nv, 9)
    T = _amps(No, Nv, "complex", 3)
    with api.CoupledCluster(No, Nv, NG) as cc:
        cc.set_inputs(g, eps)
        cc.set_amplitudes(*T)
        grad_h, S_h = cc.energy_backward(g)
        gd = _DeviceArray(g.nbytes)
        out = _DeviceArray(g.nbytes)
        try:
            gd.upload(g)
            E_d = cc.energy_forward(gd.ptr.value)
            _, S_d = cc.energy_backward(gd.ptr.value, grad=out.ptr.value)
            grad_d = np.zeros_like(g)
            out.download(grad_d)
        finally:
            gd.free()
            out.free()
        assert E_d == cc.energy_forward(g)
        assert np.array_equal(S_d, S_h)
        assert np.array_equal(grad_d, grad_h)


def test_backward_needs_amplitudes():
    No, Nv, NG = 3, 7, 5
    inp = syn.generate(No, Nv, NG, "complex", 2)
    with api.CoupledCluster(No, Nv, NG) as cc:
        cc.set_inputs(inp.gamma, inp.eps)
        with pytest.raises(api.CcsdError, match="no amplitudes"):
            cc.energy_backward()
        with pytest.raises(api.CcsdError, match="precision"):
            cc.set_amplitudes(*_amps(No, Nv, "complex", 1))
            cc.energy_backward(precision=api.PRECISION_MIXED)


def test_adjoint_at_converged_amplitudes_no32_nv256():
    """No=32, Nv=256, NG=1500 (complex, MP2 per occupied orbital as No=8,
    Nv=40): converge FP64 CCSD, then the FP64 and FP32 gradients against the
    oracle's host gradient, and S(G) against central differences of the
    forward energy (relative step 1e-5 of one G slice) at 20 seeded G, to
    1e-7 of max |S|."""
    cfg = syn.WORKLOADS["adjoint_no32_nv256_ng1500"]
    print("generating No=32 Nv=256 NG=1500", flush=True)
    inp = syn.generate_config(cfg, scale_rule="mp2_per_occupied",
                              mp2_fn=None)
    No, Nv, NG = cfg.No, cfg.Nv, cfg.NG
    with api.CoupledCluster(No, Nv, NG, field="complex",
                            maxIterations=cfg.max_iterations,
                            energyConvergence=cfg.energy_convergence,
                            amplitudesConvergence=cfg.amplitudes_convergence,
                            maxResidua=cfg.max_residua) as cc:
        cc.set_inputs(inp.gamma, inp.eps)
        res = cc.run()
        print(f"converged={res['convergenceReached']} after {res['iterations']}, "
              f"E={res['energy']['correlation']:.12f}", flush=True)
        assert res["convergenceReached"]
        E = cc.energy_forward()
        assert abs(E - res["energy"]["correlation"]) < 1e-12 * abs(E)
        grad, S = cc.energy_backward()
        grad32, S32 = cc.energy_backward(precision=api.PRECISION_FP32)
        T = cc.amplitudes()
        print("host gradient", flush=True)
        g_ref, S_ref = O.vertex_gradient(inp.gamma, T, No, "complex")
        assert rel_err(grad, g_ref) < 1e-12
        assert rel_err(S, S_ref) < 1e-12
        assert rel_err(grad32, g_ref) < 1e-5
        assert rel_err(S32, S_ref) < 1e-5
        assert abs(np.sum(S) - E) < 1e-12 * abs(E)
        rng = np.random.default_rng(2024)
        h = 1e-5
        g = np.asfortranarray(inp.gamma)
        print("finite differences", flush=True)
        for G in rng.choice(NG, size=20, replace=False):
            sl = g[G].copy()
            g[G] = sl * (1 + h)
            ep = cc.energy_forward(g)
            g[G] = sl * (1 - h)
            em = cc.energy_forward(g)
            g[G] = sl
            fd = (ep - em) / (4 * h)
            # rounding of E (~1e-16 |E|) divided by 4h bounds the difference
            assert abs(fd - S[G]) < 1e-7 * np.max(np.abs(S))
