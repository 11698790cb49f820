"""Host logic of bench.py (no GPU): the time-to-convergence cut and the
projection of a run from per-pass costs."""

import bench


def _passes(flags, seconds=1.5):
    return [{"converged": c, "seconds": seconds + 0.01 * k, "energy": -1.0 - 1e-3 * k}
            for k, c in enumerate(flags)]


def test_time_to_convergence_stops_at_the_first_converged_pass():
    # 5 warmup + 20 timed passes, the test is met at pass 15 and after
    passes = _passes([False] * 14 + [True] * 11)
    out = bench.time_to_convergence(passes)
    assert out["iterations"] == 15
    assert abs(out["seconds"] - sum(1.5 + 0.01 * k for k in range(15))) < 1e-12
    assert out["energy"] == passes[14]["energy"] and out["converged"]


def test_time_to_convergence_without_convergence_counts_every_pass():
    passes = _passes([False] * 7)
    out = bench.time_to_convergence(passes)
    assert out["iterations"] == 7 and not out["converged"]
    assert out["energy"] == passes[-1]["energy"]


def test_recorded_pass_sequences():
    seq = bench.RECORDED_PASSES_NO64_NV512
    assert len(seq["fp64"]) == 17 and set(seq["fp64"]) == {"F"}
    assert seq["mixed"].count("A") == 1 and seq["mixed"].endswith("F")
