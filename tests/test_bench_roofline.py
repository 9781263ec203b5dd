"""bench.py's roofline bookkeeping on synthetic kernel times (no GPU): the dominant category is the one with the most
time, its timer categories are the ones the timed steps bracket, and the achieved bandwidth is its algorithmic bytes
over its live average launch."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def _kt(**ms):
    kt = {f"{k}_{s}": 0 for k in ("explicit", "prep", "shock", "residual", "jacobian", "krylov", "final", "other") for s in ("ms", "n")}
    kt.update(krylov_aux_ms=0.0, krylov_idle_ms=0.0, guess_ms=0.0)
    for k, (t, n) in ms.items():
        kt[f"{k}_ms"], kt[f"{k}_n"] = t, n
    return kt


def test_dominant_category_and_its_timer_categories():
    kt = _kt(explicit=(5.0, 20), residual=(13.0, 76), jacobian=(10.5, 56), krylov=(14.4, 181), final=(5.1, 20))
    assert bench.dominant(kt) == "krylov_iter"
    assert set(bench.TIMER_CATS["krylov_iter"]) == {"krylov", "krylov_idle"}
    kt["residual_ms"] = 20.0
    assert bench.dominant(kt) == "residual" and bench.TIMER_CATS["residual"] == ("residual",)


def test_roofline_uses_the_live_measure_and_the_iterations_taken():
    n = 2048 * 1024
    full = _kt(explicit=(5.0, 20), residual=(13.0, 76), jacobian=(10.5, 56), krylov=(14.4, 181), final=(5.1, 20))
    live = _kt(krylov=(1.59, 200))   # split pairs count in krylov_n; the taken iterations come from the step stats
    res = {"kt": full, "kt_live": live, "dom": "krylov_iter", "wall": 0.0713,
           "d": {"krylov_it": 20, "stages": 20, "newton_it": 53, "steps": 10},
           "d_all": {"krylov_it": 186, "stages": 20, "newton_it": 53, "steps": 10}}
    roof, kernel_ms, frac = bench.roofline_of(res, n, True)
    avg = 1.59e-3 / 20
    assert roof["kernel"] == "krylov_iter" and abs(roof["avg_launch_ms"] - avg * 1e3) < 1e-12
    assert roof["bytes_per_launch"] == bench.BYTES_KRYLOV_FP32 * n
    assert abs(roof["achieved"] - bench.BYTES_KRYLOV_FP32 * n / avg / 1e9) < 1e-6
    assert abs(roof["frac"] - roof["achieved"] * 1e9 / bench.HBM_PEAK) < 1e-12
    assert roof["traffic"] == bench.TRAFFIC[("krylov_iter", True, 2048, 1024)]
    assert kernel_ms["krylov_iter"] == 14.4 and 0 < frac < 1
