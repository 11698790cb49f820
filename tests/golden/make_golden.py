"""Writes the golden fixtures of tests/golden from the FP64 oracle.

real_no8_nv40_ng200.json  the real No=8, Nv=40, NG=200 workload: SHA-256 of the
              generated inputs, the oracle's per-iteration (E, dE, dR) trace,
              final energies and checksums of the converged amplitudes.
small_complex.json  a small complex workload (No=6, Nv=24, NG=60) run the
              same way, plus the residuum at seeded amplitudes.
complex_no32_nv256_ng1500.json  the complex No=32, Nv=256, NG=1500 workload
              with the vertex scaled by the "mp2_per_occupied" rule (the scale
              at which the iteration converges), run the same way, plus the
              converged T1 and T2 at seeded sample positions (about 1.5
              hours of host time on 8 cores; about 35 GB of memory).
complex_no32_nv256_ng1500_residuum.json  one oracle residuum of that workload
              at seeded amplitudes: the scale, input SHA-256 values,
              checksums of R1 and R2, R1 and R2 at seeded sample positions
              (cc4s-layout flat indices) and the energy of the amplitudes.

Usage: python tests/golden/make_golden.py [real_no8_nv40_ng200|small_complex|
                                           complex_no32_nv256_ng1500|
                                           complex_no32_nv256_ng1500_residuum ...]
"""

from __future__ import annotations

import hashlib
import json
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from cc4s_amd import synthetic as syn  # noqa: E402
from oracle import ccsd_oracle as O  # noqa: E402

HERE = pathlib.Path(__file__).resolve().parent


def checksums(T):
    """Size-independent summaries of an amplitude set."""
    T1, T2 = T
    out = {}
    for name, t in (("T1", T1), ("T2", T2)):
        f = np.asfortranarray(t).ravel(order="F")
        w = np.cos(0.37 * np.arange(f.size))
        out[name] = {
            "norm": float(np.linalg.norm(f)),
            "sum_re": float(np.sum(f.real)), "sum_im": float(np.sum(f.imag)),
            "wsum_re": float(np.sum(w * f.real)),
            "wsum_im": float(np.sum(w * f.imag)),
        }
    return out


def _progress(i, e, dR):
    print(f"  iteration {i + 1}: E={np.real(e):.15f} dR={dR:.3e} "
          f"({time.strftime('%H:%M:%S')})", flush=True)


def amplitude_samples(T, seed):
    """T1 and T2 of an amplitude set at seeded positions (cc4s-layout flat
    indices), with the largest magnitude of each part."""
    out = {}
    for k, (name, t) in enumerate((("T1", T[0]), ("T2", T[1]))):
        f = np.asfortranarray(t).ravel(order="F")
        count = RESIDUUM_SAMPLES_R1 if name == "T1" else RESIDUUM_SAMPLES_R2
        pos = sample_positions(f.size, count, seed + k)
        out[name + "_samples"] = {"seed": seed + k, "positions": pos.tolist(),
                                  "re": f[pos].real.tolist(), "im": f[pos].imag.tolist()}
        out[name + "_max_abs"] = float(np.max(np.abs(f)))
    return out


def run_config(cfg, extra=None, scale_rule="max_v", samples=False):
    inp = syn.generate_config(cfg, scale_rule=scale_rule)
    P = O.CcsdProblem(inp.gamma, inp.eps, inp.No, cfg.field)
    t = time.time()
    r = O.coupled_cluster(P, max_iterations=cfg.max_iterations,
                          energy_convergence=cfg.energy_convergence,
                          amplitudes_convergence=cfg.amplitudes_convergence,
                          level_shift=cfg.level_shift,
                          max_residua=cfg.max_residua, callback=_progress)
    dt = time.time() - t
    g = np.asfortranarray(inp.gamma).tobytes(order="F")
    out = {
        "config": cfg.__dict__,
        "scale_rule": scale_rule,
        "scale": inp.scale,
        "gamma_sha256": hashlib.sha256(g).hexdigest(),
        "eps_sha256": hashlib.sha256(inp.eps.tobytes()).hexdigest(),
        "energy": r["energy"],
        "convergenceReached": r["convergenceReached"],
        "iterations": r["iterations"],
        "trace": r["trace"],
        "amplitudes": checksums(r["amplitudes"]),
        "oracle_seconds": dt,
    }
    if samples:
        out["amplitude_samples"] = amplitude_samples(r["amplitudes"], AMPLITUDE_SAMPLE_SEED)
    if extra:
        out.update(extra(P, inp))
    return out


def residuum_fixture(P, inp):
    T = syn.random_amplitudes(P.No, P.Nv, P.field, seed=77)
    R1, R2 = P.residuum(T)
    E, D, X = O.energy(P, T)
    return {"residuum_at_seeded_amplitudes": {
        "amplitude_seed": 77,
        "R1": checksums((R1, R1))["T1"], "R2": checksums((R2, R2))["T2"],
        "energy": [E.real, E.imag, D.real, D.imag, X.real, X.imag]}}


# samples of the config-size residuum fixture
RESIDUUM_AMPLITUDE_SEED = 77
RESIDUUM_SAMPLE_SEED = 2026
RESIDUUM_SAMPLES_R1, RESIDUUM_SAMPLES_R2 = 256, 2048
AMPLITUDE_SAMPLE_SEED = 3030


def sample_positions(n, count, seed):
    """Sorted distinct flat positions in [0, n), from the SplitMix64 stream
    (uniforms of stream 5), so the test draws the same positions."""
    u = syn.uniforms(seed, 5, np.arange(4 * count))
    draws = (u * n).astype(np.int64)
    _, first = np.unique(draws, return_index=True)   # first draw of each value
    pos = np.sort(draws[np.sort(first)[:count]])
    assert len(pos) == count
    return pos


def residuum_at_config_size(cfg, scale_rule="mp2_per_occupied"):
    inp = syn.generate_config(cfg, scale_rule=scale_rule)
    P = O.CcsdProblem(inp.gamma, inp.eps, inp.No, cfg.field)
    T = syn.random_amplitudes(cfg.No, cfg.Nv, cfg.field, seed=RESIDUUM_AMPLITUDE_SEED)
    t = time.time()
    R1, R2 = P.residuum(T)
    dt = time.time() - t
    E, D, X = O.energy(P, T)
    f1 = np.asfortranarray(R1).ravel(order="F")
    f2 = np.asfortranarray(R2).ravel(order="F")
    p1 = sample_positions(f1.size, RESIDUUM_SAMPLES_R1, RESIDUUM_SAMPLE_SEED)
    p2 = sample_positions(f2.size, RESIDUUM_SAMPLES_R2, RESIDUUM_SAMPLE_SEED + 1)
    g = np.asfortranarray(inp.gamma).tobytes(order="F")
    return {
        "config": cfg.__dict__,
        "scale_rule": scale_rule,
        "scale": inp.scale,
        "gamma_sha256": hashlib.sha256(g).hexdigest(),
        "eps_sha256": hashlib.sha256(inp.eps.tobytes()).hexdigest(),
        "amplitudes": {"generator": "synthetic.random_amplitudes", "seed":
                       RESIDUUM_AMPLITUDE_SEED, "t1_scale": 0.02, "t2_scale": 0.01},
        "energy": [E.real, E.imag, D.real, D.imag, X.real, X.imag],
        "R1": checksums((R1, R1))["T1"], "R2": checksums((R2, R2))["T2"],
        "R1_max_abs": float(np.max(np.abs(f1))), "R2_max_abs": float(np.max(np.abs(f2))),
        "R1_samples": {"seed": RESIDUUM_SAMPLE_SEED, "positions": p1.tolist(),
                       "re": f1[p1].real.tolist(), "im": f1[p1].imag.tolist()},
        "R2_samples": {"seed": RESIDUUM_SAMPLE_SEED + 1, "positions": p2.tolist(),
                       "re": f2[p2].real.tolist(), "im": f2[p2].imag.tolist()},
        "oracle_seconds": dt,
    }


def main(which):
    if "real_no8_nv40_ng200" in which:
        data = run_config(syn.WORKLOADS["real_no8_nv40_ng200"], residuum_fixture)
        (HERE / "real_no8_nv40_ng200.json").write_text(json.dumps(data, indent=1))
        print("real_no8_nv40_ng200", data["energy"], data["iterations"],
              data["oracle_seconds"])
    if "small_complex" in which:
        cfg = syn.CcsdConfig("complex No=6 Nv=24 NG=60", 6, 24, 60, "complex",
                             seed=42)
        data = run_config(cfg, residuum_fixture)
        (HERE / "small_complex.json").write_text(json.dumps(data, indent=1))
        print("small_complex", data["energy"], data["iterations"],
              data["oracle_seconds"])
    if "complex_no32_nv256_ng1500" in which:
        data = run_config(syn.WORKLOADS["complex_no32_nv256_ng1500"],
                          scale_rule="mp2_per_occupied", samples=True)
        (HERE / "complex_no32_nv256_ng1500.json").write_text(json.dumps(data, indent=1))
        print("complex_no32_nv256_ng1500", data["energy"], data["iterations"],
              data["oracle_seconds"])
    if "complex_no32_nv256_ng1500_residuum" in which:
        data = residuum_at_config_size(syn.WORKLOADS["complex_no32_nv256_ng1500"])
        (HERE / "complex_no32_nv256_ng1500_residuum.json").write_text(
            json.dumps(data, indent=1))
        print("complex_no32_nv256_ng1500_residuum", data["energy"][0],
              data["oracle_seconds"])


if __name__ == "__main__":
    main(sys.argv[1:] or ["real_no8_nv40_ng200", "small_complex"])
