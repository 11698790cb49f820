"""Benchmark: seconds per CCSD iteration on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 1]
                [--no-cpu-baseline]

A step is one pass of the CCSD iteration loop (residuum, denominator update,
DIIS, energy) on the seeded synthetic workload of the chosen configuration,
with the Coulomb vertex already resident in HBM.  The first W passes are
untimed (pass 1 has no residuum contraction: R = V).  Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from cc4s_amd import api  # noqa: E402
from cc4s_amd import synthetic as syn  # noqa: E402

FP64_MFMA_PEAK_TFLOPS = 78.6      # MI355X dense FP64 matrix
HBM_PEAK_GBS = 8000.0
METRIC = ("s per CCSD iteration + % MFMA peak, mixed vs FP64 GPU at "
          "|ΔE|≤1 µEh")
VARIANTS = [f"gett_{t}_{a}_{b}" for t in ("c64", "f64") for a in ("ak", "am")
            for b in ("bk", "bn")]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(cfg, inp, iteration_flops, seconds_budget=20.0):
    """Times the oracle's particle-particle ladder statements
    (oracle/ccsd_oracle.py, Ccsd.cxx:588-590) on slice pairs of the same
    workload until ~seconds_budget of host time, and converts the sustained
    rate into seconds per iteration with the iteration's algorithmic flops."""
    try:
        from threadpoolctl import threadpool_limits
    except Exception:  # pragma: no cover
        threadpool_limits = None
    from oracle import ccsd_oracle as O
    cores = min(16, os.cpu_count() or 1)
    ctx = threadpool_limits(limits=cores) if threadpool_limits else None
    try:
        No, Nv = cfg.No, cfg.Nv
        s = No
        gpp = np.ascontiguousarray(inp.gamma[:, No:, No:])
        cT = np.conj(gpp).transpose(0, 2, 1)
        T = syn.random_amplitudes(No, Nv, cfg.field, seed=5, t1_scale=0.01,
                                  t2_scale=0.01)
        X = np.ascontiguousarray(T[1])
        cmul = 8.0 if cfg.field == "complex" else 2.0
        flops = 0.0
        pairs = 0
        t0 = time.perf_counter()
        ns = (Nv + s - 1) // s
        for m in range(ns):
            for n in range(m, ns):
                a, b = n * s, m * s
                V = O.ein("Gxc,Gyd->xycd", cT[:, a:a + s, :], gpp[:, b:b + s, :])
                O.ein("xycd,cdij->xyij", V, X)
                flops += cmul * (cfg.NG * s * s * Nv * Nv + s * s * Nv * Nv * No * No)
                pairs += 1
                if time.perf_counter() - t0 > seconds_budget:
                    break
            if time.perf_counter() - t0 > seconds_budget:
                break
        dt = time.perf_counter() - t0
    finally:
        if ctx is not None:
            ctx.__exit__(None, None, None)
    rate = flops / dt
    return {
        "value": iteration_flops / rate,
        "unit": "s/iteration",
        "cores": cores,
        "kind": "port",
        "sample": (f"oracle (numpy/OpenBLAS FP64) particle-particle ladder, "
                   f"{pairs} slice pairs of size {s} of this workload in "
                   f"{dt:.1f} s = {rate / 1e9:.1f} GFLOP/s, extrapolated to "
                   f"the iteration's {iteration_flops:.3e} algorithmic flops"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; "
                          "--gpus must be 1"}))
        sys.exit(2)
    cfg = syn.CONFIGS[args.config]
    log(f"generating inputs: {cfg.name}")
    t = time.perf_counter()
    inp = syn.generate_config(cfg)
    log(f"  {time.perf_counter() - t:.1f} s, scale c = {inp.scale:.6e}")

    cc = api.CoupledCluster(cfg.No, cfg.Nv, cfg.NG, field=cfg.field,
                            maxIterations=cfg.max_iterations,
                            energyConvergence=cfg.energy_convergence,
                            amplitudesConvergence=cfg.amplitudes_convergence,
                            maxResidua=cfg.max_residua)
    t = time.perf_counter()
    cc.set_inputs(inp.gamma, inp.eps)
    log(f"inputs resident, V slices formed: {time.perf_counter() - t:.2f} s")
    for w in range(args.warmup):
        r = cc.iterate()
        log(f"warmup {r['iteration']}: E={r['energy']:.12f} dR={r['residuum_norm']:.3e} "
            f"{r['seconds']:.3f} s")
    cc.set_profiling(True)
    reps = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        r = cc.iterate()
        reps.append(r)
        log(f"step {r['iteration']}: E={r['energy']:.12f} dE={r['delta_energy']:.3e} "
            f"dR={r['residuum_norm']:.3e} {r['seconds']:.3f} s "
            f"{r['flops'] / r['seconds'] / 1e12:.2f} TFLOP/s")
    elapsed = time.perf_counter() - t0
    s_per_iter = elapsed / args.steps
    iter_flops = float(np.mean([r["flops"] for r in reps]))

    stats = {v: cc.kernel_stats(v) for v in VARIANTS}
    dom = max(stats, key=lambda v: stats[v]["ms"])
    d = stats[dom]
    avg_ms = d["ms"] / max(1, d["launches"])
    achieved = d["flops"] / (d["ms"] * 1e-3) / 1e12 if d["ms"] > 0 else 0.0
    lad_ms, lad_flops, lad_calls = cc.ladder_stats()
    log(f"kernel stats: {json.dumps(stats)}")
    log(f"ladder: {lad_calls} calls, {lad_ms:.1f} ms, "
        f"{lad_flops / (lad_ms * 1e-3) / 1e12 if lad_ms else 0:.2f} TFLOP/s")
    cc.close()

    out = {
        "metric": METRIC,
        "value": s_per_iter,
        "unit": "s/iteration",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * s_per_iter,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64/Box-Muller vertex, cc4s layout)",
        "config": {
            "workload": f"FP64 CCSD iteration, {cfg.name}, DIIS {cfg.max_residua}",
            "No": cfg.No, "Nv": cfg.Nv, "NG": cfg.NG, "field": cfg.field,
            "precision": "fp64", "parallelism": "1 GPU",
            "flops_per_iteration": iter_flops,
            "tflops_per_s": iter_flops / s_per_iter / 1e12,
            "frac_fp64_mfma_peak_iteration": iter_flops / s_per_iter / 1e12
            / FP64_MFMA_PEAK_TFLOPS,
            "ladder_tflops_per_s": (lad_flops / (lad_ms * 1e-3) / 1e12
                                    if lad_ms else None),
            "energy_last_step": reps[-1]["energy"],
        },
        "roofline": {
            "kernel": dom,
            "bound": "mfma",
            "achieved": achieved,
            "peak": FP64_MFMA_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_MFMA_PEAK_TFLOPS,
            "traffic": None,
            "launches": d["launches"],
            "avg_launch_ms": avg_ms,
            "flops_per_launch": d["flops"] / max(1, d["launches"]),
        },
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle sample) ...")
        out["cpu_baseline"] = cpu_baseline(cfg, inp, iter_flops)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
