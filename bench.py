"""Benchmark: seconds per CCSD iteration on one MI355X, and mixed vs FP64
time to convergence.

python bench.py [--gpus 1] [--steps K] [--warmup W]
                [--workload complex_no32_nv256_ng1500 | --config I]
                [--scale mp2_per_occupied|max_v] [--converge both|fp64|mixed|none]
                [--mixed-switch 1e-5,1e-7] [--no-cpu-baseline] [--adjoint auto|on|off]

A step is one pass of the CCSD iteration loop (residuum, denominator update,
DIIS, energy) on the seeded synthetic workload of the chosen configuration,
with the Coulomb vertex already resident in HBM.  The FP64 context does W
untimed passes (pass 1 has no residuum contraction: R = V), then K timed
passes: `value` is their mean wall time.  With --converge the FP64 context
keeps iterating to the convergence criteria, and one further context per
mixed switch value and refinement mode runs the mixed-precision path (split
16-bit contractions until dR < switch, an FP64 anchor pass, split passes
corrected by its defect; refinement 1 lets an FP64 pass declare convergence,
3 the corrected pass that meets the criteria) to the same criteria; every
time to convergence, its ratio to the FP64 one and |E_mixed - E_FP64| are
reported.  The first switch value is the workload's written 1e-5, the
second 1e-6.  At the FP64 fixed point of the
No=32, Nv=256, NG=1500 workload the energy adjoint (gradient with respect to the vertex and the
structure factor S(G), FP64 and FP32) is timed and checked against central
differences of the forward energy.  Prints one JSON line.

Input scale: "mp2_per_occupied" (default) scales the vertex so that the MP2
energy per occupied orbital equals that of the real No=8, Nv=40, NG=200
workload; "max_v" is max|V_abij| = 0.05 Eh, for which the iteration diverges
at larger No, Nv.  --config I selects the I-th workload of
synthetic.WORKLOADS (0 real No=8 Nv=40 NG=200, 1 complex No=32 Nv=256
NG=1500, 2 complex No=64 Nv=512 NG=3000, 3 the adjoint workload, 4 complex
No=64 Nv=640 NG=4000).
Per-pass cost does not depend on the scale.
"""

from __future__ import annotations

import argparse
import ctypes
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
METRIC = ("s per CCSD iteration + % MFMA peak, mixed vs FP64 GPU at "
          "|ΔE|≤1 µEh")
VARIANTS = [f"gett_{t}_{a}_{b}{k}" for t in ("c64", "f64") for a in ("ak", "am")
            for b in ("bk", "bn") for k in ("", "_shortk")]


def measured_traffic(kernel, cfg):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE and WRITE_SIZE, calibrated against a
    plain copy kernel), or None."""
    import re
    name = f"profiles/{kernel}_hbm_traffic_no{cfg.No}_nv{cfg.Nv}.md"
    path = os.path.join(ROOT, name)
    if not os.path.exists(path):
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


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
        done = False
        for m in range(ns):
            for n in range(m, ns):
                a, b = n * s, m * s
                V = O.ein("Gxc,Gyd->xycd", cT[:, a:a + s, :], gpp[:, b:b + s, :])
                O.ein("xycd,cdij->xyij", V, X)
                flops += cmul * (cfg.NG * s * s * Nv * Nv + s * s * Nv * Nv * No * No)
                pairs += 1
                done = time.perf_counter() - t0 > seconds_budget
                if done:
                    break
            if done:
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
        "cpu_model": cpu_model(),
        "sample": (f"oracle (numpy/OpenBLAS FP64) particle-particle ladder, "
                   f"{pairs} slice pairs of size {s} of this workload in "
                   f"{dt:.1f} s = {rate / 1e9:.1f} GFLOP/s, extrapolated to "
                   f"the iteration's {iteration_flops:.3e} algorithmic flops"),
    }


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline_full(cfg, inp):
    """Times one complete iteration of the oracle on the host: its
    CoupledCluster loop (oracle/ccsd_oracle.py, CoupledCluster.cxx:129-175)
    run for two passes, the first of which has no residuum contraction
    (R = V, Ccsd.cxx:404-409); the second pass is the residuum of every
    statement of Ccsd.cxx:381-633, the denominator update, the DIIS step and
    the energy.  Takes minutes at No=32, Nv=256."""
    try:
        from threadpoolctl import threadpool_limits
    except Exception:  # pragma: no cover
        threadpool_limits = None
    from oracle import ccsd_oracle as O
    cores = min(16, os.cpu_count() or 1)
    ctx = threadpool_limits(limits=cores) if threadpool_limits else None
    stamps = []
    try:
        P = O.CcsdProblem(inp.gamma, inp.eps, cfg.No, cfg.field)
        t0 = time.perf_counter()
        O.coupled_cluster(P, max_iterations=2, energy_convergence=0.0,
                          amplitudes_convergence=0.0, max_residua=cfg.max_residua,
                          callback=lambda *a: stamps.append(time.perf_counter()))
    finally:
        if ctx is not None:
            ctx.__exit__(None, None, None)
    first, second = stamps[0] - t0, stamps[1] - stamps[0]
    return {
        "value": second,
        "unit": "s/iteration",
        "cores": cores,
        "kind": "port",
        "cpu_model": cpu_model(),
        "sample": (f"one complete pass of the oracle's CCSD loop (numpy/OpenBLAS FP64: "
                   f"residuum, update, DIIS, energy) on this workload, timed after its "
                   f"first pass (R = V, {first:.1f} s)"),
    }


MIXED_OPTIONS = {}
# the particle-particle ladder of the timed contexts (--ladder)
LADDER = {"name": "paired"}

# Pass sequences of full runs of the complex No=64, Nv=512, NG=3000 workload
# (scale rule mp2_per_occupied, DIIS 8, convergence |dE| < 1e-10 and
# dR < 1e-8) to convergence on one MI355X with the paired ladder: FP64
# converges at pass 17; the mixed run (half split, 2048-term FP32 sums,
# switch 1e-5, refinement 1) does 8 split passes (the first is the R = V
# pass), the FP64 anchor, 5 corrected split passes and one FP64 pass.
# F = FP64, S = split, A = anchor (an FP64 and a split residuum),
# C = corrected split.
# dense half MFMA peak of MI355X and the algorithmic ceiling of a split
# multiply with 12 half products per complex MAC (8 algorithmic flops)
HALF_MFMA_PEAK_TFLOPS = 2500.0
SPLIT12_CEILING_TFLOPS = HALF_MFMA_PEAK_TFLOPS * 8.0 / 24.0
RECORDED_PASSES_NO64_NV512 = {"fp64": "F" * 17, "mixed": "S" * 8 + "A" + "C" * 5 + "F"}


def dominant_kernel(stats):
    """The contraction-kernel variant with the largest summed duration, its
    algorithmic rate and fraction of the FP64 matrix peak."""
    dom = max(stats, key=lambda v: stats[v]["ms"])
    d = stats[dom]
    achieved = d["flops"] / (d["ms"] * 1e-3) / 1e12 if d["ms"] > 0 else 0.0
    return {
        "kernel": dom, "achieved": achieved, "peak": FP64_MFMA_PEAK_TFLOPS,
        "unit": "TFLOP/s", "frac": achieved / FP64_MFMA_PEAK_TFLOPS,
        "launches": d["launches"], "avg_launch_ms": d["ms"] / max(1, d["launches"]),
        "flops_per_launch": d["flops"] / max(1, d["launches"]),
    }


def large_workload_report(wname, scale, fused_ladder=False):
    """One FP64 pass and one split (mixed) pass of a large workload, timed in
    this run after the passes of the main workload: the first pass of each
    context (R = V, no residuum contraction) is untimed, the second evaluates
    the full residuum.  Reports s/iteration, the iteration's and the
    particle-particle ladder's algorithmic rate against the FP64 matrix peak,
    the dominant kernel, the split pass's speed-up, and the time to
    convergence of both paths projected from these per-pass costs and the
    pass sequences of recorded full runs (RECORDED_PASSES_NO64_NV512)."""
    cfg = syn.WORKLOADS[wname]
    log(f"large workload: generating inputs: {cfg.name}, scale rule {scale}")
    t = time.perf_counter()
    inp = syn.generate_config(cfg, scale_rule=scale, mp2_fn=syn.mp2_on_device)
    log(f"  {time.perf_counter() - t:.1f} s, scale c = {inp.scale:.6e}")
    out = {"workload": f"{cfg.name}, DIIS {cfg.max_residua}", "No": cfg.No,
           "Nv": cfg.Nv, "NG": cfg.NG, "scale_rule": scale, "scale": inp.scale}
    with make_context(cfg, api.PRECISION_FP64, 1e-5) as cc:
        cc.set_inputs(inp.gamma, inp.eps)
        first = cc.iterate()
        cc.set_profiling(True)
        r = cc.iterate()
        stats = {v: cc.kernel_stats(v) for v in VARIANTS}
        lad_ms, lad_flops, _ = cc.ladder_stats()
        log(f"large workload FP64 pass: E={r['energy']:.12f} {r['seconds']:.3f} s "
            f"{r['flops'] / r['seconds'] / 1e12:.2f} TFLOP/s")
    ladder = lad_flops / (lad_ms * 1e-3) / 1e12 if lad_ms else None
    out["fp64"] = {
        "seconds_per_iteration": r["seconds"],
        "ladder_ms": lad_ms,
        "first_pass_seconds": first["seconds"],
        "residuum_seconds": r["residuum_seconds"],
        "update_and_diis_seconds": r["mixer_seconds"],
        "flops_per_iteration": r["flops"],
        "tflops_per_s": r["flops"] / r["seconds"] / 1e12,
        "frac_fp64_mfma_peak_iteration": r["flops"] / r["seconds"] / 1e12
        / FP64_MFMA_PEAK_TFLOPS,
        "ladder_tflops_per_s": ladder,
        "ladder_frac_fp64_mfma_peak": ladder / FP64_MFMA_PEAK_TFLOPS if ladder else None,
        "dominant_kernel": dominant_kernel(stats),
    }
    if fused_ladder:
        with make_context(cfg, api.PRECISION_FP64, 1e-5, ladder="fused") as cc:
            cc.set_inputs(inp.gamma, inp.eps)
            cc.iterate()
            cc.set_profiling(True)
            fr = cc.iterate()
            f_ms, f_flops, _ = cc.ladder_stats()
        f_rate = f_flops / (f_ms * 1e-3) / 1e12 if f_ms else None
        out["fp64_fused_ladder"] = {
            "seconds_per_iteration": fr["seconds"],
            "ladder_ms": f_ms, "ladder_tflops_per_s": f_rate,
            "ladder_frac_fp64_mfma_peak": f_rate / FP64_MFMA_PEAK_TFLOPS if f_rate else None,
            "abs_energy_difference_vs_slab": abs(fr["energy"] - r["energy"]),
        }
        log(f"large workload FP64 pass, fused ladder: {json.dumps(out['fp64_fused_ladder'])}")
    with make_context(cfg, api.PRECISION_MIXED, 1e-5) as cc:
        cc.set_inputs(inp.gamma, inp.eps)
        cc.iterate()
        m1 = cc.iterate()   # makes the tau operand's scale and packing plans
        cc.set_profiling(True)
        m = cc.iterate()
        split = cc.kernel_stats("mixed_c64")
        log(f"large workload split passes: {m1['seconds']:.3f} s, {m['seconds']:.3f} s, "
            f"E={m['energy']:.12f}")
    out["mixed"] = {
        "seconds_per_iteration": m["seconds"],
        "first_split_pass_seconds": m1["seconds"],
        "split": {0: "bf16", 1: "half"}[cc.params.mixed_split],
        "max_k": cc.params.mixed_max_k,
        "algorithmic_tflops_per_s": r["flops"] / m["seconds"] / 1e12,
        "speedup_vs_fp64_pass": r["seconds"] / m["seconds"],
    }
    # the split contractions of the timed pass (HIP events around each one:
    # operand packing inside their K loop included, scale reductions not)
    if split["ms"] > 0:
        a = split["flops"] / (split["ms"] * 1e-3) / 1e12
        out["mixed"]["split_roofline"] = {
            "kernel": "bgemm (12 half products per complex MAC)", "bound": "mfma",
            "achieved": a, "unit": "TFLOP/s", "launches": split["launches"],
            "seconds": split["ms"] * 1e-3,
            "peak": SPLIT12_CEILING_TFLOPS, "frac": a / SPLIT12_CEILING_TFLOPS,
            "frac_of_half_peak_executed": a * 3.0 / HALF_MFMA_PEAK_TFLOPS,
        }
    if wname == "complex_no64_nv512_ng3000":
        cost = {"F": r["seconds"], "S": m["seconds"], "C": m["seconds"],
                "A": r["seconds"] + m["seconds"]}

        def projected(seq):
            # the first pass of a run is the R = V pass; the first split pass
            # also makes the scale and packing plans
            t = first["seconds"] + sum(cost[k] for k in seq[1:])
            return t + (m1["seconds"] - m["seconds"]) * (seq[1] == "S")
        t64 = projected(RECORDED_PASSES_NO64_NV512["fp64"])
        tmx = projected(RECORDED_PASSES_NO64_NV512["mixed"])
        out["time_to_convergence_projected"] = {
            "fp64_seconds": t64, "mixed_seconds": tmx, "speedup": t64 / tmx,
            "pass_sequences": RECORDED_PASSES_NO64_NV512,
        }
    return out


def ladder_comparison(cfg, inp, main_passes, lad_main, steps=2):
    """The particle-particle ladder of FP64 passes in each ladder mode: the
    main loop's mode (timed there) and the others -- the V slab, the fused
    kernel (V formed in LDS, no slab) and the paired contraction (over
    c <= d, i <= j combinations of tau) -- each in a context that runs its
    first pass (R = V, no ladder) untimed and `steps` passes with the ladder
    timed by HIP events around it.  Their energies are compared with the
    main context's passes of the same index (same amplitudes up to
    rounding)."""
    def rate(ms, flops, calls):
        t = flops / (ms * 1e-3) / 1e12 if ms else None
        return {"ladder_ms_per_pass": ms / max(1, calls), "ladder_tflops_per_s": t,
                "ladder_frac_fp64_mfma_peak": t / FP64_MFMA_PEAK_TFLOPS if t else None,
                "flops_per_pass": flops / max(1, calls)}
    main = LADDER["name"]
    out = {main: rate(*lad_main)}
    diff = 0.0
    for other in api.LADDERS:
        if other == main:
            continue
        with make_context(cfg, api.PRECISION_FP64, 1e-5, ladder=other) as cc:
            cc.set_inputs(inp.gamma, inp.eps)
            passes = [cc.iterate()]
            cc.set_profiling(True)
            for _ in range(steps):
                passes.append(cc.iterate())
            out[other] = rate(*cc.ladder_stats())
            out[other]["seconds_per_iteration"] = float(
                np.mean([p["seconds"] for p in passes[1:]]))
        n = min(len(passes), len(main_passes))
        diff = max(diff, max(abs(a["energy"] - b["energy"])
                             for a, b in zip(passes[:n], main_passes[:n])))
    out["main"] = main
    out["max_abs_energy_difference_vs_main"] = diff
    for other in api.LADDERS:
        if other != main:
            out[f"{other}_over_{main}_time"] = (out[other]["ladder_ms_per_pass"]
                                               / out[main]["ladder_ms_per_pass"])
    log(f"ladder comparison: {json.dumps(out)}")
    return out


def make_context(cfg, precision, switch, restart=None, keep=None, refine=None,
                 ladder=None):
    """A context of the workload; mixed contexts take the split settings of
    the command line (MIXED_OPTIONS), every context its ladder (LADDER)
    unless one is named."""
    extra = dict(MIXED_OPTIONS) if precision == api.PRECISION_MIXED else {}
    extra["ladder"] = ladder or LADDER["name"]
    if refine is not None and precision == api.PRECISION_MIXED:
        extra["mixedRefine"] = refine
    return api.CoupledCluster(cfg.No, cfg.Nv, cfg.NG, field=cfg.field,
                              maxIterations=cfg.max_iterations,
                              energyConvergence=cfg.energy_convergence,
                              amplitudesConvergence=cfg.amplitudes_convergence,
                              maxResidua=cfg.max_residua, precision=precision,
                              mixedSwitch=switch, mixedRestart=restart,
                              mixedRestartKeep=keep, **extra)


def time_to_convergence(passes):
    """Passes, wall time and energy up to and including the first pass that
    meets the convergence test (all passes if none does): passes timed
    beyond convergence do not count."""
    n = next((k + 1 for k, p in enumerate(passes) if p["converged"]), len(passes))
    cut = passes[:n]
    return {
        "iterations": n,
        "seconds": float(sum(p["seconds"] for p in cut)),
        "energy": cut[-1]["energy"] if cut else None,
        "converged": bool(cut and cut[-1]["converged"]),
    }


def converge(cc, label, first_passes):
    """Iterates until a pass meets the convergence test or maxIterations;
    returns every pass (those given first included)."""
    passes = list(first_passes)
    while not any(p["converged"] for p in passes) and \
            len(passes) < cc.params.max_iterations:
        r = cc.iterate()
        passes.append(r)
        kind = ("fp64 ", "mixed", "anchr", "corr ")[r["pass_kind"]]
        log(f"{label} {r['iteration']:3d} {kind} "
            f"E={r['energy']:.12f} dE={r['delta_energy']:.3e} "
            f"dR={r['residuum_norm']:.3e} {r['seconds']:.3f} s")
    return passes


def device_memory():
    """(free, total) GB of the current device, from the HIP runtime the
    library loaded."""
    hip = ctypes.CDLL("libamdhip64.so.7")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
        return None
    return round(free.value / 1e9, 2), round(total.value / 1e9, 2)


PLACEMENTS = {0: "hbm", 1: "amplitude slots in host memory",
              2: "every slot in host memory"}


def phase_report(passes):
    """Mean wall time per pass of the residuum and of the amplitude update +
    DIIS step, and where the DIIS history lives."""
    if not passes:
        return None
    return {
        "passes": len(passes),
        "seconds_per_iteration": float(np.mean([p["seconds"] for p in passes])),
        "residuum_seconds": float(np.mean([p["residuum_seconds"] for p in passes])),
        "update_and_diis_seconds": float(np.mean([p["mixer_seconds"] for p in passes])),
        "diis_history": PLACEMENTS[passes[-1]["diis_placement"]],
    }


def adjoint_report(cc, gamma, energy, n_fd=20, h=1e-5, seed=2024):
    """Energy adjoint at the converged amplitudes: forward energy, FP64 and
    FP32 backward (gradient + structure factor S(G)), sum_G S(G) against E,
    FP32 against FP64, and central differences of the forward energy along
    a scaling of one G slice by 1 +- h at `n_fd` seeded G (E(1+t) =
    E_rest + (1+t)^2 S(G)), relative to max |S|."""
    t = time.perf_counter()
    e_fwd = cc.energy_forward()
    t_fwd = time.perf_counter() - t
    t = time.perf_counter()
    grad, S = cc.energy_backward()
    t64 = time.perf_counter() - t
    t = time.perf_counter()
    grad32, S32 = cc.energy_backward(precision=api.PRECISION_FP32)
    t32 = time.perf_counter() - t
    smax = float(np.max(np.abs(S)))
    rng = np.random.default_rng(seed)
    g = np.asfortranarray(gamma)
    fd_err = 0.0
    for G in rng.choice(g.shape[0], size=min(n_fd, g.shape[0]), replace=False):
        sl = g[G].copy()
        g[G] = sl * (1 + h)
        ep = cc.energy_forward(g)
        g[G] = sl * (1 - h)
        em = cc.energy_forward(g)
        g[G] = sl
        fd_err = max(fd_err, abs((ep - em) / (4 * h) - S[G]) / smax)
    scale = float(np.max(np.abs(grad)))
    out = {
        "energy_forward": e_fwd,
        "forward_vs_iteration_energy": abs(e_fwd - energy),
        "sum_S_minus_E_rel": abs(float(np.sum(S)) - e_fwd) / abs(e_fwd),
        "fp32_grad_rel_err": float(np.max(np.abs(grad32 - grad))) / scale,
        "fp32_S_rel_err": float(np.max(np.abs(S32 - S))) / smax,
        "fd_S_rel_err_max": fd_err,
        "fd_points": int(min(n_fd, g.shape[0])), "fd_step": h,
        "seconds_forward": t_fwd, "seconds_backward_fp64": t64,
        "seconds_backward_fp32": t32,
    }
    log(f"adjoint: {json.dumps(out)}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="complex_no32_nv256_ng1500",
                    choices=tuple(syn.WORKLOADS))
    ap.add_argument("--config", type=int, default=None,
                    help="index into the workload list (overrides --workload)")
    ap.add_argument("--scale", default="mp2_per_occupied",
                    choices=("mp2_per_occupied", "max_v"))
    ap.add_argument("--converge", default="both",
                    choices=("both", "fp64", "mixed", "none"))
    ap.add_argument("--mixed-switch", default="1e-5,1e-6",
                    help="comma-separated dR thresholds of the mixed -> FP64 "
                         "switch; one mixed run per value (the first is the "
                         "workload's written 1e-5)")
    ap.add_argument("--mixed-restart", default="auto", choices=("auto", "on", "off", "both"),
                    help="DIIS history at the mixed -> FP64 switch (auto: restart "
                         "when the switch comes at dR < 2e-6)")
    ap.add_argument("--mixed-restart-keep", default="0",
                    help="comma-separated numbers of recent DIIS pairs a restart keeps")
    ap.add_argument("--mixed-refine", default="1,3",
                    help="comma-separated refinement modes after the split passes "
                         "(0 FP64 passes, 1/2 FP64 anchor + defect-corrected split "
                         "passes + FP64 passes, 3 the same, a corrected pass may "
                         "declare convergence)")
    ap.add_argument("--mixed-split", default=None, choices=("bf16", "half"),
                    help="split format of the mixed passes (default: the library's)")
    ap.add_argument("--mixed-max-k", type=int, default=None,
                    help="longest FP32 sum of a split contraction (default: the library's)")
    ap.add_argument("--mixed-steps", type=int, default=0,
                    help="also time this many mixed-precision passes (after "
                         "--warmup untimed ones) without converging")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-baseline-full", action="store_true",
                    help="time one complete oracle iteration on the host instead of "
                         "the bounded ladder sample (minutes at No=32, Nv=256)")
    ap.add_argument("--large", default="complex_no64_nv512_ng3000",
                    choices=("none",) + tuple(syn.WORKLOADS),
                    help="after the main workload, time one FP64 and one split "
                         "pass of this workload (none: skip)")
    ap.add_argument("--ladder", default="paired", choices=tuple(api.LADDERS),
                    help="particle-particle ladder of the timed contexts")
    ap.add_argument("--ladder-compare", default="auto", choices=("auto", "on", "off"),
                    help="time FP64 passes with the other ladder modes next to the "
                         "main one (auto: on for the No=32 Nv=256 workload)")
    ap.add_argument("--large-fused-ladder", action="store_true",
                    help="also time an FP64 pass of the large workload with the fused ladder")
    ap.add_argument("--triples", action="store_true",
                    help="time the (T) correction at the FP64 fixed point")
    ap.add_argument("--adjoint", default="auto", choices=("auto", "on", "off"),
                    help="energy adjoint at the FP64 fixed point (auto: on "
                         "for the No=32, Nv=256, NG=1500 workloads)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the CCSD iteration runs on one GPU per process "
                          "(no sharding across GPUs); --gpus must be 1"}))
        sys.exit(2)
    switches = [float(x) for x in str(args.mixed_switch).split(",") if x]
    LADDER["name"] = args.ladder
    if args.mixed_split is not None:
        MIXED_OPTIONS["mixedSplit"] = args.mixed_split
    if args.mixed_max_k is not None:
        MIXED_OPTIONS["mixedMaxK"] = args.mixed_max_k
    refines = ([int(x) for x in str(args.mixed_refine).split(",") if x]
               if args.mixed_refine is not None else [None])
    if args.steps < 1 and args.converge != "none":
        ap.error("--steps must be at least 1 unless --converge none")
    wname = (list(syn.WORKLOADS)[args.config] if args.config is not None
             else args.workload)
    cfg = syn.WORKLOADS[wname]
    log(f"generating inputs: {cfg.name}, scale rule {args.scale}")
    t = time.perf_counter()
    big = cfg.No * cfg.Nv > 16384
    inp = syn.generate_config(cfg, scale_rule=args.scale,
                              mp2_fn=syn.mp2_on_device if big else None)
    log(f"  {time.perf_counter() - t:.1f} s, scale c = {inp.scale:.6e}")

    out_conv = {}
    adjoint = None
    triples = None
    reps = []
    stats = {}
    lad = (0.0, 0.0, 0)
    value = None
    iter_flops = None
    if args.converge in ("both", "fp64") or (args.converge == "none" and args.steps > 0):
        cc = make_context(cfg, api.PRECISION_FP64, switches[0])
        t = time.perf_counter()
        cc.set_inputs(inp.gamma, inp.eps)
        log(f"inputs resident, V slices formed: {time.perf_counter() - t:.2f} s; "
            f"device memory free/total after setup: {device_memory()}")
        passes = []
        for w in range(args.warmup):
            r = cc.iterate()
            passes.append(r)
            log(f"warmup {r['iteration']}: E={r['energy']:.12f} "
                f"dR={r['residuum_norm']:.3e} {r['seconds']:.3f} s")
        cc.set_profiling(True)
        t0 = time.perf_counter()
        for k in range(args.steps):
            r = cc.iterate()
            reps.append(r)
            passes.append(r)
            log(f"step {r['iteration']}: E={r['energy']:.12f} dE={r['delta_energy']:.3e} "
                f"dR={r['residuum_norm']:.3e} {r['seconds']:.3f} s "
                f"{r['flops'] / r['seconds'] / 1e12:.2f} TFLOP/s")
        value = (time.perf_counter() - t0) / args.steps
        iter_flops = float(np.mean([r["flops"] for r in reps]))
        phases = phase_report(reps)
        stats = {v: cc.kernel_stats(v) for v in VARIANTS}
        lad = cc.ladder_stats()
        cc.set_profiling(False)
        main_passes = list(passes)
        if args.converge in ("both", "fp64"):
            passes = converge(cc, "fp64", passes)
            out_conv["fp64"] = time_to_convergence(passes)
            want = args.adjoint == "on" or (args.adjoint == "auto"
                                            and wname in ("complex_no32_nv256_ng1500",
                                                          "adjoint_no32_nv256_ng1500"))
            if want and out_conv["fp64"]["converged"]:
                adjoint = adjoint_report(cc, inp.gamma, passes[-1]["energy"])
            if args.triples:
                t = time.perf_counter()
                e_t = cc.triples()
                dt = time.perf_counter() - t
                # six particle (Nv^4) and six hole (No Nv^3) contractions per
                # hole triple i <= j <= k: No (No+1) (No+2) (Nv^4 + No Nv^3)
                # MACs, 8 flop per complex MAC
                o, v = cfg.No, cfg.Nv
                mac = o * (o + 1) * (o + 2) * (v ** 4 + o * v ** 3)
                fl = (8.0 if cfg.field == "complex" else 2.0) * mac
                triples = {"energy": e_t, "seconds": dt,
                           "algorithmic_flops": fl, "tflops_per_s": fl / dt / 1e12}
                log(f"(T): {json.dumps(triples)}")
        cc.close()
    mixed_iter = None
    if args.mixed_steps > 0:
        cc = make_context(cfg, api.PRECISION_MIXED, switches[0])
        cc.set_inputs(inp.gamma, inp.eps)
        log(f"mixed context: device memory free/total after setup: {device_memory()}")
        for w in range(args.warmup):
            r = cc.iterate()
            log(f"mixed warmup {r['iteration']}: E={r['energy']:.12f} "
                f"dR={r['residuum_norm']:.3e} {r['seconds']:.3f} s")
        mreps = []
        for k in range(args.mixed_steps):
            r = cc.iterate()
            mreps.append(r)
            log(f"mixed step {r['iteration']}: E={r['energy']:.12f} "
                f"dR={r['residuum_norm']:.3e} {r['seconds']:.3f} s "
                f"({'mixed' if r['precision_used'] else 'fp64'})")
        cc.close()
        mixed_iter = phase_report([r for r in mreps if r["precision_used"]])
        if mixed_iter and value is not None:
            mixed_iter["speedup_vs_fp64_iteration"] = value / mixed_iter["seconds_per_iteration"]
            mixed_iter["algorithmic_tflops_per_s"] = (
                iter_flops / mixed_iter["seconds_per_iteration"] / 1e12)
    restarts = {"auto": ["auto"], "on": [True], "off": [False],
                "both": [True, False]}[args.mixed_restart]
    keeps = [int(x) for x in str(args.mixed_restart_keep).split(",") if x]
    runs = [(sw, rs, kp, rf) for sw in switches for rs in restarts for kp in keeps
            for rf in refines]
    for n, (sw, rs, kp, rf) in enumerate(runs if args.converge in ("both", "mixed") else []):
        key = "mixed" if n == 0 else f"mixed_switch_{sw:g}" + (
            "" if rs in (True, "auto") else "_norestart") + (f"_keep{kp}" if kp else "") + (
            f"_refine{rf}" if rf else "")
        cc = make_context(cfg, api.PRECISION_MIXED, sw, rs, kp, rf)
        cc.set_inputs(inp.gamma, inp.eps)
        passes = converge(cc, key, [])
        ttc = time_to_convergence(passes)
        passes = passes[:ttc["iterations"]]
        mixed_times = [p["seconds"] for p in passes[1:] if p["precision_used"]]
        fp64_times = [p["seconds"] for p in passes[1:] if not p["precision_used"]]
        out_conv[key] = {
            "iterations": ttc["iterations"],
            "mixed_iterations": sum(1 for p in passes if p["precision_used"]),
            "fp64_passes": len(fp64_times),
            "seconds": ttc["seconds"],
            "seconds_per_mixed_iteration": (float(np.mean(mixed_times))
                                            if mixed_times else None),
            "seconds_per_fp64_pass": (float(np.mean(fp64_times))
                                      if fp64_times else None),
            "energy": ttc["energy"],
            "converged": ttc["converged"],
            "switch": sw,
            "diis_restart_at_switch": rs,
            "diis_pairs_kept": kp,
            "refine": cc.params.mixed_refine,
            "split": {0: "bf16", 1: "half"}[cc.params.mixed_split],
            "max_k": cc.params.mixed_max_k,
            "pass_kinds": "".join("FSAC"[p["pass_kind"]] for p in passes),
        }
        cc.close()
        if "fp64" in out_conv:
            out_conv[key]["speedup_time_to_convergence"] = (
                out_conv["fp64"]["seconds"] / out_conv[key]["seconds"])
            out_conv[key]["abs_energy_difference"] = abs(
                out_conv["fp64"]["energy"] - out_conv[key]["energy"])

    ladders = None
    if value is not None and cfg.field == "complex" and (
            args.ladder_compare == "on" or (args.ladder_compare == "auto"
                                            and wname == "complex_no32_nv256_ng1500")):
        ladders = ladder_comparison(cfg, inp, main_passes, lad)

    large = None
    if args.large != "none" and args.large != wname:
        large = large_workload_report(args.large, args.scale, args.large_fused_ladder)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "s/iteration",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * value if value is not None else None,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": (f"synthetic (seeded SplitMix64/Box-Muller vertex, cc4s layout, "
                 f"scale rule {args.scale})"),
        "config": {
            "workload": f"FP64 CCSD iteration, {cfg.name}, DIIS {cfg.max_residua}",
            "No": cfg.No, "Nv": cfg.Nv, "NG": cfg.NG, "field": cfg.field,
            "precision": "fp64", "parallelism": "1 GPU",
            "time_to_convergence": out_conv,
        },
    }
    if value is not None:
        out["config"]["fp64_iteration_phases"] = phases
    if mixed_iter is not None:
        out["config"]["mixed_iteration"] = mixed_iter
    if adjoint is not None:
        out["config"]["adjoint"] = adjoint
    if triples is not None:
        out["config"]["triples"] = triples
    if large is not None:
        out["config"]["large_workload"] = large
    if ladders is not None:
        out["config"]["ladder_comparison"] = ladders
    if value is not None:
        dk = dominant_kernel(stats)
        lad_ms, lad_flops, _ = lad
        out["config"].update({
            "flops_per_iteration": iter_flops,
            "tflops_per_s": iter_flops / value / 1e12,
            "frac_fp64_mfma_peak_iteration": iter_flops / value / 1e12
            / FP64_MFMA_PEAK_TFLOPS,
            "ladder_tflops_per_s": (lad_flops / (lad_ms * 1e-3) / 1e12
                                    if lad_ms else None),
        })
        # the complex FP64 kernels run the three-product complex multiply
        # (6 real flops of MFMA work per complex MAC; generation 2: four
        # products, 8); `achieved` credits 8 per complex MAC
        gen = int(os.environ.get("CC4S_GETT_GENERATION", "3"))
        executed = (6.0 if gen >= 3 else 8.0) / 8.0 if cfg.field == "complex" else 1.0
        out["roofline"] = {
            "kernel": dk["kernel"],
            "bound": "mfma",
            "achieved": dk["achieved"],
            "peak": FP64_MFMA_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": dk["frac"],
            "traffic": measured_traffic(dk["kernel"], cfg),
            "frac_executed_mfma": dk["frac"] * executed,
            "launches": dk["launches"],
            "avg_launch_ms": dk["avg_launch_ms"],
            "flops_per_launch": dk["flops_per_launch"],
        }
        log(f"kernel stats: {json.dumps(stats)}")
        if args.cpu_baseline_full:
            log("cpu baseline (one complete oracle iteration) ...")
            out["cpu_baseline"] = cpu_baseline_full(cfg, inp)
        elif not args.no_cpu_baseline:
            log("cpu baseline (oracle sample) ...")
            out["cpu_baseline"] = cpu_baseline(cfg, inp, iter_flops)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
