"""bench.py's roofline bookkeeping on synthetic kernel times (no GPU): the dominant category is the one with the most
time (the shock budget included), its timer categories are the ones the timed steps bracket, and the achieved rate is
its algorithmic bytes (HBM-bound categories) or FP64 operations (the shock budget) over its live average launch."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def _kt(**ms):
    kt = {f"{k}_{s}": 0 for k in ("explicit", "prep", "shock", "residual", "jacobian", "krylov", "final", "other") for s in ("ms", "n")}
    kt.update(krylov_aux_ms=0.0, krylov_idle_ms=0.0, guess_ms=0.0, shock_rk_steps=0)
    for k, v in ms.items():
        if k == "shock_rk_steps":
            kt[k] = v
            continue
        kt[f"{k}_ms"], kt[f"{k}_n"] = v
    return kt


def test_dominant_category_and_its_timer_categories():
    kt = _kt(explicit=(5.0, 20), residual=(13.0, 76), jacobian=(10.5, 56), krylov=(14.4, 181), final=(5.1, 20))
    assert bench.dominant(kt) == "krylov_iter"
    assert set(bench.TIMER_CATS["krylov_iter"]) == {"krylov", "krylov_idle"}
    kt["residual_ms"] = 20.0
    assert bench.dominant(kt) == "residual" and bench.TIMER_CATS["residual"] == ("residual",)
    kt["shock_ms"] = 30.0
    assert bench.dominant(kt) == "shock" and bench.TIMER_CATS["shock"] == ("shock",)


def test_hbm_roofline_uses_the_live_measure_and_the_iterations_taken():
    n = 2048 * 1024
    full = _kt(explicit=(5.0, 20), residual=(13.0, 76), jacobian=(10.5, 56), krylov=(14.4, 181), final=(5.1, 20))
    live = _kt(krylov=(1.59, 200))   # split pairs count in krylov_n; the taken iterations come from the step stats
    res = {"kt": full, "kt_live": live, "dom": "krylov_iter", "wall": 0.0713,
           "d": {"krylov_it": 20, "stages": 20, "newton_it": 53, "steps": 10},
           "d_all": {"krylov_it": 186, "stages": 20, "newton_it": 53, "steps": 10}, "fused": 1}
    roof, kernel_ms, frac = bench.roofline_of(res, n, True)
    avg = 1.59e-3 / 20
    assert roof["kernel"] == "krylov_iter" and roof["bound"] == "hbm" and abs(roof["avg_launch_ms"] - avg * 1e3) < 1e-12
    assert roof["bytes_per_launch"] == bench.BYTES_KRYLOV_FP32 * n == 148 * n
    assert abs(roof["achieved"] - bench.BYTES_KRYLOV_FP32 * n / avg / 1e9) < 1e-6
    assert abs(roof["frac"] - roof["achieved"] * 1e9 / bench.HBM_PEAK) < 1e-12
    # the red-black data flow's own count is reported beside the contract one, never in place of it
    rb = roof["by_kernel"]["krylov_iter_rb172"] / roof["by_kernel"]["krylov_iter"]
    assert abs(rb - 172 / 148) < 1e-3
    assert kernel_ms["krylov_iter"] == 14.4 and 0 < frac < 1


import pytest  # noqa: E402


@pytest.mark.parametrize("fused,ops", [(0, 555), (1, 456)])
def test_shock_roofline_counts_fp64_operations_per_rk_step(fused, ops):
    n = 4096 * 4096
    full = _kt(explicit=(4.0, 2), shock=(50.0, 2), residual=(9.0, 7), shock_rk_steps=600_000_000)
    live = _kt(shock=(100.0, 4), shock_rk_steps=1_300_000_000)
    res = {"kt": full, "kt_live": live, "dom": "shock", "wall": 0.5,
           "d": {"krylov_it": 10, "stages": 8, "newton_it": 20, "steps": 4},
           "d_all": {"krylov_it": 5, "stages": 2, "newton_it": 5, "steps": 1}, "fused": fused}
    roof, _, _ = bench.roofline_of(res, n, True)
    assert roof["bound"] == "fp64" and roof["kernel"] == "shock" and roof["unit"] == "TFLOP/s"
    assert bench.SHOCK_RK_FLOPS[fused] == ops
    achieved = 1_300_000_000 / 4 * ops / (100.0e-3 / 4)
    assert abs(roof["achieved"] - achieved / 1e12) < 1e-9 and abs(roof["frac"] - achieved / bench.FP64_PEAK) < 1e-12
    assert abs(roof["by_kernel"]["shock_fp64"] - 600_000_000 * ops / 0.05 / bench.FP64_PEAK) < 1e-4
