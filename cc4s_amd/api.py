"""ctypes binding of the C ABI in include/cc4s_ccsd.h.

Arrays cross the boundary in the cc4s layout (first index fastest), i.e.
numpy arrays in Fortran order indexed in the reference's index order
(T2[a,b,i,j], Gamma[G,p,q]).  The library must be built
(``python -m cc4s_amd.build``); there is no fallback path.
"""

from __future__ import annotations

import ctypes as C
import math
import pathlib

import numpy as np

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libcc4s_ccsd.so"

FIELD_REAL = 0
FIELD_COMPLEX = 1
PRECISION_FP64 = 0
PRECISION_MIXED = 1
PRECISION_FP32 = 2
METHODS = {"Ccsd": 0, "Drccd": 1}

# every symbol declared in include/cc4s_ccsd.h
EXPORTS = (
    "cc4s_ccsd_default_params", "cc4s_ccsd_create", "cc4s_ccsd_set_inputs",
    "cc4s_ccsd_set_inputs_device", "cc4s_ccsd_set_amplitudes",
    "cc4s_ccsd_iterate", "cc4s_ccsd_run", "cc4s_ccsd_get_amplitudes",
    "cc4s_ccsd_residuum", "cc4s_ccsd_energy",
    "cc4s_ccsd_energy_grad_forward", "cc4s_ccsd_energy_grad_backward",
    "cc4s_ccsd_triples",
    "cc4s_ccsd_transition_structure_factor", "cc4s_finite_size_correction",
    "cc4s_ccsd_set_profiling",
    "cc4s_ccsd_ladder_stats", "cc4s_ccsd_kernel_stats",
    "cc4s_ccsd_last_error", "cc4s_ccsd_destroy",
    "cc4s_contract", "cc4s_ccsd_energy_noalloc",
)


class Params(C.Structure):
    _fields_ = [
        ("field", C.c_int),
        ("No", C.c_int64), ("Nv", C.c_int64), ("NG", C.c_int64),
        ("max_iterations", C.c_int64),
        ("energy_convergence", C.c_double),
        ("amplitudes_convergence", C.c_double),
        ("level_shift", C.c_double),
        ("max_residua", C.c_int64),
        ("ppl", C.c_int),
        ("integrals_slice_size", C.c_int64),
        ("precision", C.c_int),
        ("mixed_switch", C.c_double),
        ("device", C.c_int),
        ("method", C.c_int),
        ("linearized", C.c_int),
        ("adjacent_pairs_exchange", C.c_int),
        ("mixed_restart", C.c_int),
        ("mixed_restart_keep", C.c_int),
        ("mixed_refine", C.c_int),
        ("mixed_split", C.c_int),
        ("mixed_max_k", C.c_int64),
        ("ladder", C.c_int),
    ]


class Iteration(C.Structure):
    _fields_ = [
        ("iteration", C.c_int64),
        ("energy", C.c_double), ("delta_energy", C.c_double),
        ("residuum_norm", C.c_double),
        ("direct", C.c_double), ("exchange", C.c_double),
        ("seconds", C.c_double), ("flops", C.c_double),
        ("converged", C.c_int), ("precision_used", C.c_int),
        ("diis_placement", C.c_int),
        ("residuum_seconds", C.c_double), ("mixer_seconds", C.c_double),
        ("pass_kind", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Result(C.Structure):
    _fields_ = [
        ("correlation", C.c_double), ("direct", C.c_double),
        ("exchange", C.c_double), ("second_order", C.c_double),
        ("unit", C.c_double), ("convergence_reached", C.c_int),
        ("iterations", C.c_int64),
    ]


class CcsdError(RuntimeError):
    pass


_lib = None


def load_library(path: pathlib.Path | str | None = None):
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = pathlib.Path(path) if path else LIB_PATH
    if not p.exists():
        raise CcsdError(f"{p} not built: run python -m cc4s_amd.build")
    lib = C.CDLL(str(p))
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    ctxp = C.c_void_p
    lib.cc4s_ccsd_default_params.argtypes = [C.POINTER(Params)]
    lib.cc4s_ccsd_default_params.restype = None
    lib.cc4s_ccsd_create.argtypes = [C.POINTER(Params), C.POINTER(ctxp)]
    lib.cc4s_ccsd_set_inputs.argtypes = [ctxp, vp, vp]
    lib.cc4s_ccsd_set_inputs_device.argtypes = [ctxp, vp, vp]
    lib.cc4s_ccsd_set_amplitudes.argtypes = [ctxp, vp, vp]
    lib.cc4s_ccsd_iterate.argtypes = [ctxp, C.POINTER(Iteration)]
    lib.cc4s_ccsd_run.argtypes = [ctxp, C.POINTER(Result), C.POINTER(Iteration),
                                  C.c_int64]
    lib.cc4s_ccsd_get_amplitudes.argtypes = [ctxp, vp, vp]
    lib.cc4s_ccsd_residuum.argtypes = [ctxp, vp, vp, vp, vp]
    lib.cc4s_ccsd_energy.argtypes = [ctxp, vp, vp, dp]
    lib.cc4s_ccsd_energy_grad_forward.argtypes = [ctxp, vp, C.c_int, dp]
    lib.cc4s_ccsd_triples.argtypes = [ctxp, dp]
    lib.cc4s_ccsd_transition_structure_factor.argtypes = [ctxp, vp, C.c_int64, vp, vp]
    lib.cc4s_finite_size_correction.argtypes = [vp, C.c_int64, vp, vp, C.c_int64,
                                                C.c_double, C.c_double, C.c_double, dp]
    lib.cc4s_ccsd_energy_grad_backward.argtypes = [ctxp, vp, C.c_int, vp, C.c_int,
                                                   vp, C.c_int]
    lib.cc4s_ccsd_set_profiling.argtypes = [ctxp, C.c_int]
    lib.cc4s_ccsd_ladder_stats.argtypes = [ctxp, dp, dp, C.POINTER(C.c_int64)]
    lib.cc4s_ccsd_kernel_stats.argtypes = [ctxp, C.c_char_p, C.POINTER(C.c_int64),
                                           dp, dp]
    lib.cc4s_ccsd_last_error.argtypes = [ctxp]
    lib.cc4s_ccsd_last_error.restype = C.c_char_p
    lib.cc4s_ccsd_destroy.argtypes = [ctxp]
    lib.cc4s_ccsd_destroy.restype = None
    i64p = C.POINTER(C.c_int64)
    lib.cc4s_contract.argtypes = [C.c_int, C.c_double, vp, C.c_char_p, i64p, C.c_int,
                                  vp, C.c_char_p, i64p, C.c_int, C.c_double, vp,
                                  C.c_char_p, i64p, C.c_int, C.c_int]
    lib.cc4s_ccsd_energy_noalloc.argtypes = [C.c_int, C.c_int64, C.c_int64, vp, vp,
                                             vp, dp]
    for name in ("cc4s_ccsd_create", "cc4s_ccsd_set_inputs",
                 "cc4s_ccsd_set_inputs_device", "cc4s_ccsd_set_amplitudes",
                 "cc4s_ccsd_iterate", "cc4s_ccsd_run", "cc4s_ccsd_get_amplitudes",
                 "cc4s_ccsd_residuum", "cc4s_ccsd_energy",
                 "cc4s_ccsd_set_profiling", "cc4s_ccsd_ladder_stats",
                 "cc4s_ccsd_kernel_stats", "cc4s_contract",
                 "cc4s_ccsd_energy_noalloc"):
        getattr(lib, name).restype = C.c_int
    if path is None:
        _lib = lib
    return lib


def default_params() -> Params:
    p = Params()
    load_library().cc4s_ccsd_default_params(C.byref(p))
    return p


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def _fortran(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


MIXED_SPLITS = {"bf16": 0, "half": 1}
# particle-particle ladder of complex passes: V slab through HBM, one fused
# kernel that forms V in LDS (FP64 passes), or the slab contraction over pairs
# of tau's symmetric and antisymmetric combinations (half the multiply-adds)
LADDERS = {"slab": 0, "fused": 1, "paired": 2}


def finite_size_correction(grid, B, S, N=20, vertex_unit=1.0, grid_unit=1.0,
                           potential_unit=1.0):
    """(corrected, uncorrected) of FiniteSizeCorrection's interpolation:
    grid [NG, 3] Cartesian momenta, B [3, 3] reciprocal lattice vectors (rows
    Gi, Gj, Gk), S [NG] the structure factor.  Host computation."""
    lib = load_library()
    g = np.ascontiguousarray(grid, dtype=np.float64)
    b = np.ascontiguousarray(B, dtype=np.float64)
    s = np.ascontiguousarray(S, dtype=np.float64)
    if g.ndim != 2 or g.shape[1] != 3 or b.shape != (3, 3) or s.shape != (g.shape[0],):
        raise ValueError("expecting grid [NG, 3], B [3, 3], S [NG]")
    out = (C.c_double * 2)()
    rc = lib.cc4s_finite_size_correction(_ptr(g), g.shape[0], _ptr(b), _ptr(s), int(N),
                                         float(vertex_unit), float(grid_unit),
                                         float(potential_unit), out)
    if rc != 0:
        raise CcsdError(lib.cc4s_ccsd_last_error(None).decode())
    return out[0], out[1]


class CoupledCluster:
    """One CCSD run on the GPU: the CoupledCluster step with method Ccsd and
    DiisMixer, arguments named as in a cc4s operation file."""

    def __init__(self, No, Nv, NG, field="complex", maxIterations=16,
                 energyConvergence=1e-7, amplitudesConvergence=1e-6,
                 levelShift=0.0, maxResidua=4, ppl=True,
                 integralsSliceSize=0, device=0, precision=PRECISION_FP64,
                 mixedSwitch=1e-5, method="Ccsd", linearized=False,
                 adjacentPairsExchange=False, mixedRestart=None, mixedRestartKeep=None,
                 mixedRefine=None, mixedSplit=None, mixedMaxK=None, ladder="paired"):
        self.lib = load_library()
        p = default_params()
        p.field = FIELD_COMPLEX if field == "complex" else FIELD_REAL
        p.No, p.Nv, p.NG = No, Nv, NG
        p.max_iterations = maxIterations
        p.energy_convergence = energyConvergence
        p.amplitudes_convergence = amplitudesConvergence
        p.level_shift = levelShift
        p.max_residua = maxResidua
        p.ppl = 1 if ppl else 0
        p.integrals_slice_size = integralsSliceSize
        p.device = device
        p.precision = precision
        p.mixed_switch = mixedSwitch
        if method not in METHODS:
            raise CcsdError(f"Unknown method: '{method}'")
        p.method = METHODS[method]
        p.linearized = 1 if linearized else 0
        p.adjacent_pairs_exchange = 1 if adjacentPairsExchange else 0
        if mixedRestart is not None:
            p.mixed_restart = {False: 0, True: 1, "auto": 2}[mixedRestart]
        if mixedRestartKeep is not None:
            p.mixed_restart_keep = int(mixedRestartKeep)
        if mixedRefine is not None:
            p.mixed_refine = int(mixedRefine)
        if mixedSplit is not None:
            if mixedSplit not in MIXED_SPLITS:
                raise CcsdError(f"unknown split format {mixedSplit!r}")
            p.mixed_split = MIXED_SPLITS[mixedSplit]
        if mixedMaxK is not None:
            p.mixed_max_k = int(mixedMaxK)
        if ladder not in LADDERS:
            raise CcsdError(f"unknown ladder {ladder!r}")
        p.ladder = LADDERS[ladder]
        self.params = p
        self.field = field
        self.No, self.Nv, self.NG = No, Nv, NG
        self.dtype = np.complex128 if field == "complex" else np.float64
        self._ctx = C.c_void_p()
        rc = self.lib.cc4s_ccsd_create(C.byref(p), C.byref(self._ctx))
        if rc != 0:
            raise CcsdError(self.lib.cc4s_ccsd_last_error(None).decode())

    def _check(self, rc):
        if rc != 0:
            raise CcsdError(self.lib.cc4s_ccsd_last_error(self._ctx).decode())

    def close(self):
        if self._ctx:
            self.lib.cc4s_ccsd_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_inputs(self, gamma, eps):
        g = _fortran(gamma, np.complex128)
        Np = self.No + self.Nv
        if g.shape != (self.NG, Np, Np):
            raise ValueError(f"gamma shape {g.shape} != {(self.NG, Np, Np)}")
        e = self._eps(eps)
        self._check(self.lib.cc4s_ccsd_set_inputs(self._ctx, _ptr(g), _ptr(e)))

    def set_inputs_device(self, gamma_device_ptr: int, eps):
        e = self._eps(eps)
        self._check(self.lib.cc4s_ccsd_set_inputs_device(
            self._ctx, C.c_void_p(gamma_device_ptr), _ptr(e)))

    def _eps(self, eps):
        e = np.ascontiguousarray(eps, dtype=np.float64)
        if e.shape != (self.No + self.Nv,):
            raise ValueError(f"eps shape {e.shape} != {(self.No + self.Nv,)}")
        return e

    def _amplitude_arrays(self, T1, T2):
        """(T1, T2) as Fortran arrays of the context's field, shapes checked:
        the library reads Nv*No and Nv*Nv*No*No values."""
        t1, t2 = _fortran(T1, self.dtype), _fortran(T2, self.dtype)
        if t1.shape != (self.Nv, self.No):
            raise ValueError(f"T1 shape {t1.shape} != {(self.Nv, self.No)}")
        if t2.shape != (self.Nv, self.Nv, self.No, self.No):
            raise ValueError(f"T2 shape {t2.shape} != "
                             f"{(self.Nv, self.Nv, self.No, self.No)}")
        return t1, t2

    def set_amplitudes(self, T1, T2):
        t1, t2 = self._amplitude_arrays(T1, T2)
        self._check(self.lib.cc4s_ccsd_set_amplitudes(self._ctx, _ptr(t1), _ptr(t2)))

    def iterate(self):
        it = Iteration()
        self._check(self.lib.cc4s_ccsd_iterate(self._ctx, C.byref(it)))
        return it.as_dict()

    def run(self):
        res = Result()
        cap = int(self.params.max_iterations) + 1
        trace = (Iteration * cap)()
        self._check(self.lib.cc4s_ccsd_run(self._ctx, C.byref(res), trace, cap))
        n = res.iterations
        energy = {"correlation": res.correlation, "direct": res.direct,
                  "exchange": res.exchange, "unit": res.unit}
        if not math.isnan(res.second_order):
            energy["secondOrder"] = res.second_order
        return {"energy": energy,
                "convergenceReached": bool(res.convergence_reached),
                "iterations": n,
                "trace": [trace[k].as_dict() for k in range(n)]}

    def amplitudes(self):
        T1 = np.zeros((self.Nv, self.No), dtype=self.dtype, order="F")
        T2 = np.zeros((self.Nv, self.Nv, self.No, self.No), dtype=self.dtype, order="F")
        self._check(self.lib.cc4s_ccsd_get_amplitudes(self._ctx, _ptr(T1), _ptr(T2)))
        return T1, T2

    def residuum(self, amplitudes=None):
        R1 = np.zeros((self.Nv, self.No), dtype=self.dtype, order="F")
        R2 = np.zeros((self.Nv, self.Nv, self.No, self.No), dtype=self.dtype, order="F")
        if amplitudes is None:
            t1 = t2 = None
        else:
            t1, t2 = self._amplitude_arrays(*amplitudes)
        self._check(self.lib.cc4s_ccsd_residuum(self._ctx, _ptr(t1), _ptr(t2),
                                                _ptr(R1), _ptr(R2)))
        return R1, R2

    def energy(self, amplitudes):
        t1, t2 = self._amplitude_arrays(*amplitudes)
        out = (C.c_double * 6)()
        self._check(self.lib.cc4s_ccsd_energy(self._ctx, _ptr(t1), _ptr(t2), out))
        return complex(out[0], out[1]), complex(out[2], out[3]), complex(out[4], out[5])

    def _vertex_arg(self, gamma):
        """(pointer, on_device, keep-alive) for a vertex argument: None (the
        input vertex), a device address (int), or a host array."""
        if gamma is None:
            return None, 0, None
        if isinstance(gamma, int):
            return C.c_void_p(gamma), 1, None
        g = _fortran(gamma, np.complex128)
        Np = self.No + self.Nv
        if g.shape != (self.NG, Np, Np):
            raise ValueError(f"gamma shape {g.shape} != {(self.NG, Np, Np)}")
        return _ptr(g), 0, g

    def triples(self):
        """(T) correlation energy of the current amplitudes."""
        e = C.c_double()
        self._check(self.lib.cc4s_ccsd_triples(self._ctx, C.byref(e)))
        return e.value

    def transition_structure_factor(self, V, U=None):
        """S[G] of FiniteSizeCorrection at the current amplitudes: V [NGg]
        the Coulomb potential, U [NGg, NG] the singular vectors (None: the
        vertex is on the grid)."""
        V = _fortran(V, np.float64)
        NGg = V.shape[0]
        u = None if U is None else _fortran(U, np.complex128)
        if u is not None and u.shape != (NGg, self.NG):
            raise ValueError(f"singular vectors shape {u.shape} != {(NGg, self.NG)}")
        S = np.zeros(NGg)
        self._check(self.lib.cc4s_ccsd_transition_structure_factor(
            self._ctx, None if u is None else _ptr(u), NGg, _ptr(V), _ptr(S)))
        return S

    def energy_forward(self, gamma=None):
        """Correlation energy at the current amplitudes with the vertex
        `gamma` (None: the input vertex; int: a device address)."""
        ptr, dev, _keep = self._vertex_arg(gamma)
        e = C.c_double()
        self._check(self.lib.cc4s_ccsd_energy_grad_forward(self._ctx, ptr, dev,
                                                           C.byref(e)))
        return e.value

    def energy_backward(self, gamma=None, grad=True, precision=PRECISION_FP64):
        """(grad, S): grad = dE/dRe Gamma + i dE/dIm Gamma as a host array
        [NG, Np, Np] (grad=True), written to a device address (grad=int), or
        skipped (grad=False -> None); S[G] the structure factor."""
        ptr, dev, _keep = self._vertex_arg(gamma)
        Np = self.No + self.Nv
        S = np.zeros(self.NG, dtype=np.float64)
        out = None
        if grad is True:
            out = np.zeros((self.NG, Np, Np), dtype=np.complex128, order="F")
            gptr, gdev = _ptr(out), 0
        elif isinstance(grad, int) and grad is not False:
            gptr, gdev = C.c_void_p(grad), 1
        else:
            gptr, gdev = None, 0
        self._check(self.lib.cc4s_ccsd_energy_grad_backward(
            self._ctx, ptr, dev, gptr, gdev, _ptr(S), precision))
        return out, S

    def set_profiling(self, enabled=True):
        self._check(self.lib.cc4s_ccsd_set_profiling(self._ctx, 1 if enabled else 0))

    def kernel_stats(self, variant):
        n, ms, fl = C.c_int64(), C.c_double(), C.c_double()
        self._check(self.lib.cc4s_ccsd_kernel_stats(
            self._ctx, variant.encode(), C.byref(n), C.byref(ms), C.byref(fl)))
        return {"launches": n.value, "ms": ms.value, "flops": fl.value}

    def ladder_stats(self):
        ms, fl, n = C.c_double(), C.c_double(), C.c_int64()
        self._check(self.lib.cc4s_ccsd_ladder_stats(self._ctx, C.byref(ms),
                                                    C.byref(fl), C.byref(n)))
        return ms.value, fl.value, n.value


def contract(alpha, A, idx_a, B, idx_b, beta, C_, idx_c, conj_a=False,
             conj_b=False, device=0, precision=PRECISION_FP64):
    """GPU contraction with host arrays (indexed in index-string order)."""
    lib = load_library()
    cplx = any(np.iscomplexobj(x) for x in (A, B, C_))
    dt = np.complex128 if cplx else np.float64
    a, b = _fortran(A, dt), _fortran(B, dt)
    c = _fortran(C_, dt).copy(order="F")
    la = (C.c_int64 * max(1, a.ndim))(*a.shape)
    lb = (C.c_int64 * max(1, b.ndim))(*b.shape)
    lc = (C.c_int64 * max(1, c.ndim))(*c.shape)
    rc = lib.cc4s_contract(FIELD_COMPLEX if cplx else FIELD_REAL, alpha,
                           _ptr(a), idx_a.encode(), la, int(conj_a),
                           _ptr(b), idx_b.encode(), lb, int(conj_b),
                           beta, _ptr(c), idx_c.encode(), lc, precision, device)
    if rc != 0:
        raise CcsdError(lib.cc4s_ccsd_last_error(None).decode())
    return c


def energy_noalloc(field, No, Nv, T_device, V_device, stream=None):
    """CoupledCluster::getEnergy of device-resident buffers through the
    allocation-free entry point: T_device is the address of one amplitude set
    [T1; T2] and V_device that of V_abij, both in the cc4s layout on the
    current device.  Returns (E, D, X) as complex numbers."""
    lib = load_library()
    f = {"real": FIELD_REAL, "complex": FIELD_COMPLEX}.get(field, field)
    out = (C.c_double * 6)()
    rc = lib.cc4s_ccsd_energy_noalloc(int(f), int(No), int(Nv), C.c_void_p(T_device),
                                      C.c_void_p(V_device), C.c_void_p(stream), out)
    if rc != 0:
        raise CcsdError(lib.cc4s_ccsd_last_error(None).decode())
    return complex(out[0], out[1]), complex(out[2], out[3]), complex(out[4], out[5])
