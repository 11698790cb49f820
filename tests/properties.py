"""Properties of the CCSD equations that hold at any size, evaluated in
slices of the last hole index so that arrays of tens of GB need only
slice-sized temporaries.  Arrays are in the cc4s layout: numpy arrays in
Fortran order indexed T2[a,b,i,j], Gamma[G,p,q].

Orbital phase gauge: multiplying every spin orbital by a phase, phi_p ->
u_p phi_p with |u_p| = 1, changes the vertex to
Gamma^G_pq -> conj(u_p) Gamma^G_pq u_q.  The integrals the step forms,
V_abij = sum_G conj(Gamma^G_ia) Gamma^G_bj (VertexCoulombIntegrals.cxx:
143-171), pick up conj(u_a) conj(u_b) u_i u_j; so do the amplitudes and the
residuum of every closed-shell coupled-cluster equation (each index is summed
once as an orbital and once as its conjugate), while the energy, a sum of
conj(V) times amplitudes, does not change.  A dropped conjugation, a swapped
index or a transposed operand in any term breaks this.
"""

from __future__ import annotations

import numpy as np


def phases(Np: int, seed: int) -> np.ndarray:
    """u_p = exp(i theta_p), theta_p uniform in [0, 2 pi), seeded."""
    theta = np.random.default_rng(seed).uniform(0.0, 2.0 * np.pi, Np)
    return np.exp(1j * theta)


def gauge_vertex(gamma: np.ndarray, u: np.ndarray) -> None:
    """Gamma^G_pq <- Gamma^G_pq (conj(u_p) u_q), in place, one q at a time.
    The products are written out in real arithmetic (separate multiplies,
    no fused multiply-add), so the factor and the result of (q, p) are
    exactly the conjugates of those of (p, q): a Hermitian vertex stays
    exactly Hermitian and its diagonal exactly real."""
    c, s = u.real, u.imag
    for q in range(gamma.shape[2]):
        fr = c * c[q] + s * s[q]           # conj(u_p) u_q
        fi = c * s[q] - s * c[q]
        blk = gamma[:, :, q]
        gr, gi = blk.real.copy(), blk.imag.copy()
        blk.real[...] = gr * fr[None, :] - gi * fi[None, :]
        blk.imag[...] = gr * fi[None, :] + gi * fr[None, :]


def amplitude_phases(u: np.ndarray, No: int) -> np.ndarray:
    """phi_ai = conj(u_a) u_i, [Nv, No] (virtual a = No + a in the vertex)."""
    return np.conj(u[No:])[:, None] * u[None, :No]


def gauge_amplitudes(T1, T2, phi) -> None:
    """T1_ai *= phi_ai, T2_abij *= phi_ai phi_bj, in place."""
    T1 *= phi
    for j in range(T2.shape[3]):
        T2[:, :, :, j] *= phi[:, None, :] * phi[None, :, j][:, :, None]


def max_abs(x) -> float:
    if x.ndim < 4:
        return float(np.max(np.abs(x)))
    return max(float(np.max(np.abs(x[..., j]))) for j in range(x.shape[-1]))


def max_abs_diff(x, y) -> float:
    if x.ndim < 4:
        return float(np.max(np.abs(x - y)))
    return max(float(np.max(np.abs(x[..., j] - y[..., j])))
               for j in range(x.shape[-1]))


def rel_diff(x, y) -> float:
    """max |x - y| / max |y|, sliced."""
    return max_abs_diff(x, y) / max_abs(y)


def pair_asymmetry(R2) -> float:
    """max |R_abij - R_baji| / max |R|: every closed-shell doubles residuum
    and amplitude set satisfies R_abij = R_baji."""
    worst = 0.0
    for j in range(R2.shape[3]):
        d = R2[:, :, :, j] - R2[:, :, j, :].transpose(1, 0, 2)
        worst = max(worst, float(np.max(np.abs(d))))
    return worst / max_abs(R2)


def gauge_defect(R2g, R2, phi) -> float:
    """max |R'_abij - phi_ai phi_bj R_abij| / max |R|."""
    worst = 0.0
    for j in range(R2.shape[3]):
        f = phi[:, None, :] * phi[None, :, j][:, :, None]
        worst = max(worst, float(np.max(np.abs(R2g[:, :, :, j] - f * R2[:, :, :, j]))))
    return worst / max_abs(R2)


def mp2_closed_form(V, eps, No) -> float:
    """Second-order energy of canonical closed-shell orbitals from the pphh
    integrals (SecondOrderPerturbationTheory.cxx:108-127):
    sum_abij (-2 |V_abij|^2 + V_abji conj V_abij) / (e_a + e_b - e_i - e_j)."""
    eh, ep = np.asarray(eps[:No]), np.asarray(eps[No:])
    dph = ep[:, None] - eh[None, :]
    e = 0.0
    for j in range(V.shape[3]):
        D = dph[:, None, :] + dph[None, :, j][:, :, None]          # [a,b,i]
        Tc = np.conj(V[:, :, :, j]) / D
        e += float(np.real(-2.0 * np.sum(V[:, :, :, j] * Tc)
                           + np.sum(V[:, :, j, :] * Tc)))
    return e
