"""The identity behind the paired particle-particle ladder (ladder_pairs.hip),
on the host: with tau_cdij = tau_dcji (the pair symmetry of T2 and of
T1 x T1), the slab contraction of Ccsd.cxx:590,

    R_xyij = sum_cd V_xycd tau_cdij,

equals Rp + Rm for i < j, Rp - Rm at (j, i) and Rp on the diagonal, with

    Rp_xy(ij) = sum_{c<=d} Vp_xy(cd) tau+_(cd)(ij),  Vp = V_cd + V_dc (V_cc),
    Rm_xy(ij) = sum_{c<d}  Vm_xy(cd) tau-_(cd)(ij),  Vm = V_cd - V_dc,
    tau+- = (tau_cdij +- tau_cdji) / 2.

The check is against the direct contraction (numpy einsum) on random
complex V of no symmetry and random pair-symmetric tau, including Nv = 1 and
No = 1; a tau without the pair symmetry must fail it (the identity uses the
symmetry, as the reference's mirrored slice-pair blocks do).  The GPU
kernels are checked against the oracle in tests/test_gpu_ladder.py.
"""

import numpy as np
import pytest


def _tri(n, strict):
    return [(c, d) for d in range(n) for c in range(d + (0 if strict else 1))]


def paired_contraction(V, tau):
    """V [c, x, d, y], tau [c, d, i, j] -> R [x, y, i, j] through the pair
    combinations (pair index c + d(d+1)/2 for c <= d, c + d(d-1)/2 for c < d,
    the same for (i, j), as ladder_pairs.hip lays them out)."""
    v, cx, _, ny = V.shape
    o = tau.shape[2]
    kp, km = _tri(v, False), _tri(v, True)
    np_, nm = _tri(o, False), _tri(o, True)
    Tp = np.array([[0.5 * (tau[c, d, i, j] + tau[c, d, j, i]) for (i, j) in np_]
                   for (c, d) in kp]).reshape(len(kp), len(np_))
    Tm = np.array([[0.5 * (tau[c, d, i, j] - tau[c, d, j, i]) for (i, j) in nm]
                   for (c, d) in km]).reshape(len(km), len(nm))
    Vp = np.array([V[c, :, d, :] + V[d, :, c, :] if c < d else V[c, :, c, :]
                   for (c, d) in kp]).reshape(len(kp), cx, ny)
    Vm = np.array([V[c, :, d, :] - V[d, :, c, :] for (c, d) in km]).reshape(len(km), cx, ny)
    Rp = np.einsum("kxy,kn->xyn", Vp, Tp)
    # no (c < d) pairs (Nv = 1): Rm is zero, as the library clears it
    Rm = (np.einsum("kxy,kn->xyn", Vm, Tm) if len(km) and len(nm)
          else np.zeros((cx, ny, len(nm)), complex))
    R = np.zeros((cx, ny, o, o), complex)
    pos_p = {p: n for n, p in enumerate(np_)}
    pos_m = {p: n for n, p in enumerate(nm)}
    for i in range(o):
        for j in range(o):
            if i == j:
                R[:, :, i, j] = Rp[:, :, pos_p[(i, i)]]
            elif i < j:
                R[:, :, i, j] = Rp[:, :, pos_p[(i, j)]] + Rm[:, :, pos_m[(i, j)]]
            else:
                R[:, :, i, j] = Rp[:, :, pos_p[(j, i)]] - Rm[:, :, pos_m[(j, i)]]
    return R


def _inputs(v, o, cx, ny, seed, symmetric=True):
    rng = np.random.default_rng(seed)
    V = rng.standard_normal((v, cx, v, ny)) + 1j * rng.standard_normal((v, cx, v, ny))
    tau = rng.standard_normal((v, v, o, o)) + 1j * rng.standard_normal((v, v, o, o))
    if symmetric:
        tau = 0.5 * (tau + tau.transpose(1, 0, 3, 2))
    return V, tau


@pytest.mark.parametrize("v,o,cx,ny", [(5, 3, 2, 3), (9, 1, 1, 1), (1, 3, 1, 1),
                                       (7, 4, 3, 2), (12, 2, 2, 5)])
def test_paired_contraction_equals_the_slab_contraction(v, o, cx, ny):
    V, tau = _inputs(v, o, cx, ny, seed=v * 31 + o)
    direct = np.einsum("cxdy,cdij->xyij", V, tau)
    R = paired_contraction(V, tau)
    assert np.abs(R - direct).max() <= 1e-13 * np.abs(direct).max()


def test_multiply_adds_are_halved():
    v, o = 256, 32
    full = v * v * o * o
    paired = (v * (v + 1) // 2) * (o * (o + 1) // 2) + (v * (v - 1) // 2) * (o * (o - 1) // 2)
    assert paired == (full + v * o) // 2


def test_identity_needs_the_pair_symmetry():
    V, tau = _inputs(6, 3, 2, 2, seed=5, symmetric=False)
    direct = np.einsum("cxdy,cdij->xyij", V, tau)
    R = paired_contraction(V, tau)
    assert np.abs(R - direct).max() > 1e-3 * np.abs(direct).max()
