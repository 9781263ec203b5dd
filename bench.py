"""Benchmark: 2D IMEX cell-updates/s of the SPB 6+1 semi-implicit time step on one MI355X.

Headline workload: 2D low-Mach cavitation, 4096 x 4096 cells on [0, 1]^2 (water gamma 4.4, p_inf 6e8 Pa, with gas
volume fraction 1e-2, gamma 1.4; p0 = 1e5 Pa; u = 100 sign(x - 0.5), v = 100 sign(y - 0.5), each with +-1% per-cell
perturbations from the seed; transmissive boundaries), FP64 state, material CFL 0.5, LSDIRK2, Newton tolerance 1e-13
and BiCGStab tolerance 1e-13 (the reference defaults), in three forms: plain (the reference's algorithm, FP64
BiCGStab with the finite-difference Jacobian), mixed (FP32 BiCGStab with FP64 defect correction, analytic Jacobian)
and learned (mixed + the learned Newton guess). A "step" is one accepted LSDIRK2 time step (2 stages with all their
Newton and BiCGStab work). --warmup steps run first and cover the start-up ramp (the first 17 steps of this workload
are first-order steps whose dt grows from the acoustic to the material CFL step), then --breakdown steps with every
kernel category timed, then exactly --steps full LSDIRK2 steps are timed; every stepper call returns after its HIP
stream synchronises, so the timed region is bracketed by device synchronisation on both sides, and the state stays
resident in HBM throughout.

Prints one JSON line with the metric (value: the learned form's cell-updates/s), the three forms' rates, their per-field
relative L1 against the plain FP64 form after the same steps, a closed-box conservation check of each form, the
roofline of the head form's dominant kernel, every category's fraction of its roofline, and the CPU baseline (the
oracle's PC_omp-path restatement on the host's physical cores, one step of the same grid resumed from the GPU state).
Legs: the 2048 x 1024 shock-bubble (a window after the shock has reached the bubble), the 4096 x 2048 perturbed
interface, and the 4096-problem batched 1D sweep.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12     # MI355X HBM3E, B/s
FP64_PEAK = 78.6e12   # MI355X FP64, FLOP/s
# algorithmic bytes per interior cell per launch (FP64, 2D): compulsory reads + writes of the reference's data flow
# (face donors recomputed from Qe, not stored): explicit stage 232, stage preparation 184, residual 256, residual +
# Jacobian 304, BiCGStab iteration 296 (FP32 vectors and matrix: 148), final state + stage derivative 376
BYTES_2D = {"explicit": 232, "prep": 184, "residual": 256, "jacobian": 304, "krylov_iter": 296, "final": 376}
BYTES_KRYLOV_FP32 = 148
# the FP32 form's own iteration (red-black sweep preconditioner, five FP32 passes), bytes per row: kb_p 40 (r, p, v,
# 1/d and the four off-diagonals in; new p and M^-1 p out), kb_v 32 (five diagonals, M^-1 p, r0 in; v out), kb_s 36
# (r, v, 1/d and the four off-diagonals in; s, M^-1 s out), kb_t 32 (five diagonals, M^-1 s, s in; t out), kb_xr 32
# (M^-1 p, M^-1 s, s, t, x, r0 in; x, r out); a black row's red neighbours come through the caches
BYTES_KRYLOV_FP32_RB = 172
# FP64 operations of one RK45 step of the shock budget's profile integration (divisions counted as one operation, an
# FMA as two). The reference's arithmetic (prof_step): six right-hand sides of 43 (three divisions per phase, v2,
# num / den, dy1, the pressure line), stage inputs 9 + 18 + 27 + 36 + 45 plus 2 each, the error and fifth-order sums
# 69, the step-size power 72 (log 36, exp 34, the division and the product), the scales 6, the end test 2 and the step
# update 3: 555. The fused form (prof_step_fused, the bench's mixed and learned forms): six right-hand sides of 33 (v2
# 3, the pressure line 4, n_k 2, x_k 4, den 5, num 6, the one reciprocal and dy0, dy1 9), stage inputs 15 products
# h B plus 45 FMAs over three components and the 5 abscissae (2 each), sums, y5 and error 66, the power 72, the end
# test 2 and the step update 3 (the scales are per cell): 456.
SHOCK_RK_FLOPS = {0: 555, 1: 456}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30, help="timed LSDIRK2 steps of each form of the headline workload")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--breakdown", type=int, default=4, help="steps with every kernel category timed, before the timed ones")
    ap.add_argument("--n", type=int, default=4096, help="headline grid: n x n cells")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: one thread per physical core available (at most 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile", action="store_true", help="short run for rocprofv3 (no CPU legs, headline only)")
    ap.add_argument("--legs", default=None,
                    help="comma list of cavitation (the headline), bubble, interface, batch; default all (cavitation "
                         "only with --profile)")
    ap.add_argument("--forms", default=None, help="comma list of plain, mixed, learned (default all three)")
    ap.add_argument("--graph", action="store_true", help="graph mode: stage heads and Krylov chunks as captured hipGraphs")
    ap.add_argument("--bubble-warmup", type=int, default=260,
                    help="warm-up steps of the 2048x1024 shock-bubble leg (the shock reaches the bubble near step 250)")
    ap.add_argument("--bubble-steps", type=int, default=20)
    ap.add_argument("--interface-warmup", type=int, default=40)
    ap.add_argument("--interface-steps", type=int, default=10)
    ap.add_argument("--batch-steps", type=int, default=8, help="timed steps of the batched 1D leg (0: skip the leg)")
    ap.add_argument("--batch-warmup", type=int, default=12)
    ap.add_argument("--batch-cpu-problems", type=int, default=64)
    return ap.parse_args()


FORM_KW = {"plain": {}, "mixed": {"krylov_fp32": 1, "analytic_jacobian": 1, "shock_fused": 1},
           "learned": {"krylov_fp32": 1, "analytic_jacobian": 1, "learned_guess": 1, "shock_fused": 1}}
STAT_KEYS = ("steps", "stages", "newton_it", "krylov_it", "linsolves", "rejects", "guess_used", "guess_fallback", "guess_skipped")


def log(msg):
    """progress on stderr (the JSON line alone goes to stdout)"""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


GRAPH = {}   # {"graph": 1} in graph mode (--graph)


def make_stepper(make, form):
    from diffusionimex_amd import spb6
    cfg, U = make(**FORM_KW[form], **GRAPH)
    st = spb6.Stepper(cfg, device=0)
    if form == "learned":
        from diffusionimex_amd import mlp
        mlp.load_into(st)
    st.set_state(U.reshape(cfg.ny, cfg.nx, 8))
    return st, cfg, U


def run_form(make, form, warmup, steps, breakdown):
    """warm up, time `breakdown` steps category by category, then `steps` steps with only the dominant category's
    launches bracketed by events; returns the result dict and the stepper"""
    from diffusionimex_amd import spb6
    st, cfg, U0 = make_stepper(make, form)
    n = cfg.nx * cfg.ny
    log(f"{form}: {cfg.nx}x{cfg.ny}, {warmup} warm-up steps")
    rc = st.run(1.0, warmup)
    if rc != 0:
        raise SystemExit(f"{form}: warm-up failed: {spb6.STATUS.get(rc, rc)}")
    s0 = st.stats()
    kt_all = None
    if breakdown > 0:
        st.enable_timing(True)
        rc = st.run(1.0, breakdown)
        if rc != 0:
            raise SystemExit(f"{form}: breakdown run failed: {spb6.STATUS.get(rc, rc)}")
        kt_all = st.kernel_times()
    s1 = st.stats()
    dom = dominant(kt_all) if kt_all else None
    if dom:
        st.enable_timing(True, categories=TIMER_CATS[dom])
    t0 = time.perf_counter()
    rc = st.run(1.0, steps)
    t1 = time.perf_counter()
    if rc != 0:
        raise SystemExit(f"{form}: timed run failed: {spb6.STATUS.get(rc, rc)}")
    s2 = st.stats()
    kt_live = st.kernel_times()
    st.enable_timing(False)
    wall = t1 - t0
    log(f"{form}: {steps} timed steps in {wall:.2f} s")
    d = {k: s2[k] - s1[k] for k in STAT_KEYS}
    d_all = {k: s1[k] - s0[k] for k in STAT_KEYS}
    res = {"value": n * d["steps"] / wall, "wall": wall, "ms_per_step": 1e3 * wall / max(d["steps"], 1),
           "newton_per_stage": d["newton_it"] / max(d["stages"], 1), "krylov_per_newton": d["krylov_it"] / max(d["linsolves"], 1),
           "krylov_per_step": d["krylov_it"] / max(d["steps"], 1), "d": d, "kt": kt_all, "d_all": d_all, "dom": dom,
           "kt_live": kt_live, "t": st.t, "n": n, "fused": int(cfg.shock_fused)}
    if form == "learned":
        res["guess"] = {k: d[k] for k in ("guess_used", "guess_fallback", "guess_skipped")}
    return res, st


# timer categories (Stepper.TIMING_CATEGORIES) behind each roofline category
TIMER_CATS = {"explicit": ("explicit",), "prep": ("prep",), "shock": ("shock",), "residual": ("residual",),
              "jacobian": ("jacobian",), "krylov_iter": ("krylov", "krylov_idle"), "final": ("final",)}


def categories(kt, krylov_it, n, fp32):
    """HBM-bound roofline category -> (timed ms, launches, algorithmic bytes per launch)"""
    kb = BYTES_KRYLOV_FP32 if fp32 else BYTES_2D["krylov_iter"]
    return {
        "explicit": (kt["explicit_ms"], kt["explicit_n"], BYTES_2D["explicit"] * n),
        "prep": (kt["prep_ms"], kt["prep_n"], BYTES_2D["prep"] * n),
        "residual": (kt["residual_ms"], kt["residual_n"], BYTES_2D["residual"] * n),
        "jacobian": (kt["jacobian_ms"], kt["jacobian_n"], BYTES_2D["jacobian"] * n),
        # per BiCGStab iteration taken: each launched iteration is timed by its own HIP events, and those launched past a
        # solve's end (whole chunks are launched; they return at once) are kept apart in krylov_idle
        "krylov_iter": (kt["krylov_ms"], krylov_it, kb * n),
        "final": (kt["final_ms"], kt["final_n"], BYTES_2D["final"] * n),
    }


def dominant(kt):
    """the category with the most kernel time (the shock budget included)"""
    t = {k: v[0] for k, v in categories(kt, 1, 1, False).items()}
    t["shock"] = kt["shock_ms"]
    return max(t, key=lambda k: t[k])


def roofline_of(res, n, fp32, traffic=None):
    """roofline of the dominant category over its live launches in the timed region, and every category's fraction
    over the breakdown steps"""
    kt, d = res["kt"], res["d"]
    cats = categories(kt, res["d_all"]["krylov_it"], n, fp32)
    kernel_ms = {k: round(v[0], 3) for k, v in cats.items()}
    for k in ("shock", "other", "krylov_aux", "krylov_idle", "guess"):
        kernel_ms[k] = round(kt[f"{k}_ms"], 3)
    by_kernel = {k: round(v[2] * max(v[1], 1) / (v[0] / 1e3) / HBM_PEAK, 4) for k, v in cats.items() if v[0] > 0}
    if fp32 and kt["krylov_ms"] > 0:
        by_kernel["krylov_iter_rb172"] = round(BYTES_KRYLOV_FP32_RB * n * res["d_all"]["krylov_it"] / (kt["krylov_ms"] / 1e3) / HBM_PEAK, 4)
    if kt["shock_ms"] > 0:
        by_kernel["shock_fp64"] = round(kt["shock_rk_steps"] * SHOCK_RK_FLOPS[res["fused"]] / (kt["shock_ms"] / 1e3) / FP64_PEAK, 4)
    dom, live = res["dom"], res["kt_live"]
    if dom == "shock":
        ms, cnt = live["shock_ms"], live["shock_n"]
        avg_s = (ms / 1e3) / max(cnt, 1)
        flops = live["shock_rk_steps"] * SHOCK_RK_FLOPS[res["fused"]] / max(cnt, 1)
        achieved = flops / avg_s if avg_s > 0 else 0.0
        roof = {"bound": "fp64", "kernel": "shock", "achieved": achieved / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK, "traffic": traffic, "flops_per_launch": flops, "rk_steps_per_launch": live["shock_rk_steps"] / max(cnt, 1),
                "avg_launch_ms": avg_s * 1e3, "launches": cnt,
                "note": "FP64 VALU issue-bound: a correctly rounded FP64 division costs about 7.4 FMA issue slots and the "
                        "profile right-hand sides are division-dominated, so the FP64 peak (FMA = 2 operations) is not reachable"}
    else:
        ms, cnt, byts = categories(live, d["krylov_it"], n, fp32)[dom]
        avg_s = (ms / 1e3) / max(cnt, 1)
        achieved = byts / avg_s if avg_s > 0 else 0.0
        roof = {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": traffic, "bytes_per_launch": byts, "avg_launch_ms": avg_s * 1e3,
                "launches": cnt}
    roof["by_kernel"] = by_kernel
    kb = BYTES_KRYLOV_FP32 if fp32 else BYTES_2D["krylov_iter"]
    step_bytes = (BYTES_2D["explicit"] + BYTES_2D["prep"] + BYTES_2D["final"]) * d["stages"] + BYTES_2D["jacobian"] * d["newton_it"] \
        + BYTES_2D["residual"] * (d["stages"] + d["newton_it"]) + kb * d["krylov_it"]
    return roof, kernel_ms, step_bytes * n / res["wall"] / HBM_PEAK


def field_l1(A, B):
    A, B = A.reshape(-1, 8), B.reshape(-1, 8)
    return [float(np.abs(A[:, k] - B[:, k]).sum() / max(np.abs(B[:, k]).sum(), 1e-300)) for k in range(8)]


def totals(U):
    """sum m1, sum m2, sum rho u, sum rho v, sum E over the interior cells (the reference's conserved_totals, PC:1104-1107,
    with the momenta)"""
    U = U.reshape(-1, 8)
    return np.array([U[:, k].sum(dtype=np.float64) for k in (1, 2, 3, 4, 7)])


def conservation_check(a, forms, n=1024, steps=30):
    """closed box: the headline's initial state on an n x n periodic grid (every flux leaves one cell and enters another)
    -- relative change of sum m1, m2, rho u, rho v, E over `steps` steps for each form"""
    from diffusionimex_amd import cases
    from diffusionimex_amd.spb6 import BC_PERIODIC
    out = {}
    for form in forms:
        st, cfg, U0 = make_stepper(lambda **kw: cases.cavitation_2d(n, n, seed=a.seed, cfl=0.5, bc=(BC_PERIODIC,) * 4, **kw), form)
        T0 = totals(U0)
        rc = st.run(1.0, steps)
        T1 = totals(st.state())
        st.close()
        scale = np.array([abs(T0[0]), abs(T0[1]), np.abs(U0.reshape(-1, 8)[:, 3]).sum(), np.abs(U0.reshape(-1, 8)[:, 4]).sum(), abs(T0[4])])
        out[form] = {"rc": rc, "rel_change": {k: float((T1[i] - T0[i]) / scale[i]) for i, k in enumerate(("m1", "m2", "rho_u", "rho_v", "E"))}}
    out["workload"] = f"{n}x{n} periodic cavitation, {steps} steps (momenta relative to sum |rho u|, sum |rho v|)"
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}); multi-GPU sharding is design-only", file=sys.stderr)
        sys.exit(2)
    _pin_cpu_threads(a)
    if a.graph:
        GRAPH["graph"] = 1
    if a.legs is None:
        a.legs = "cavitation" if a.profile else "cavitation,bubble,interface,batch"
    legs = set(x for x in a.legs.split(",") if x)
    out = run_headline(a) if "cavitation" in legs else {"legs": sorted(legs)}
    if "bubble" in legs and not a.profile:
        log("shock-bubble leg")
        out["bubble"] = run_bubble(a)
    if "interface" in legs and a.interface_steps > 0 and not a.profile:
        log("interface leg")
        out["interface"] = run_interface(a)
    if "batch" in legs and a.batch_steps > 0 and not a.profile:
        log("batched 1D leg")
        out["batch"] = run_batch(a)
    if a.graph:
        out["graph_mode"] = True
    log("done")
    print(json.dumps(out))


def _forms(a):
    from diffusionimex_amd import mlp
    forms = [f for f in (a.forms or "plain,mixed,learned").split(",") if f]
    if "learned" in forms and not os.path.exists(mlp.WEIGHTS):
        forms.remove("learned")
    return forms


def run_headline(a):
    """The 4096^2 cavitation workload in each form: the JSON contract fields, the forms' rates and errors against the
    plain FP64 form, the head form's dominant-kernel roofline and the oracle CPU baseline."""
    from diffusionimex_amd import cases
    N = a.n
    make = lambda **kw: cases.cavitation_2d(N, N, seed=a.seed, cfl=0.5, **kw)
    forms = _forms(a)
    results, states = {}, {}
    st_plain = cfg_plain = None
    for form in forms:
        res, st = run_form(make, form, a.warmup, a.steps, a.breakdown)
        states[form] = st.state()
        results[form] = res
        if form == "plain":
            st_plain, cfg_plain = st, st.cfg
        else:
            st.close()
    head = "learned" if "learned" in results else ("mixed" if "mixed" in results else "plain")
    r = results[head]
    n = N * N
    roof, kernel_ms, hbm_frac = roofline_of(r, n, head != "plain") if r["kt"] else (None, None, None)
    forms_out = {}
    for form, res in results.items():
        fo = {k: round(res[k], 4) if isinstance(res[k], float) else res[k]
              for k in ("value", "ms_per_step", "newton_per_stage", "krylov_per_newton", "krylov_per_step")}
        fo["kernel_ms_per_step"] = {k: round(v / max(res["d_all"]["steps"], 1), 3) for k, v in roofline_of(res, n, form != "plain")[1].items()} \
            if res["kt"] else None
        if "plain" in states and form != "plain":
            fo["field_l1_vs_fp64"] = field_l1(states[form], states["plain"])
            fo["max_field_l1_vs_fp64"] = max(fo["field_l1_vs_fp64"])
            fo["speedup_vs_plain"] = res["value"] / results["plain"]["value"]
        if "guess" in res:
            fo["guess"] = res["guess"]
        forms_out[form] = fo
    out = {
        "metric": "2D IMEX cell-updates/s + BiCGStab iters/step, learned+mixed vs plain FP64",
        "value": r["value"], "unit": "cell-updates/s", "n_gpus": 1, "steps": r["d"]["steps"], "warmup": a.warmup + a.breakdown,
        "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if head == "plain" else "f64 (FP32 BiCGStab inner iterations)", "data": "synthetic",
        "config": {"workload": f"2D low-Mach cavitation {N}x{N} on [0,1]^2 ({head} form; water 4.4/6e8 with gas fraction 1e-2, "
                               f"u, v = +-100 m/s diverging with 1% seeded perturbations, transmissive)",
                   "cells": n, "seed": a.seed, "cfl": 0.5, "newton_tol": 1e-13, "krylov_tol": 1e-13, "form": head,
                   "timed_steps": f"{a.warmup + a.breakdown + 1}..{a.warmup + a.breakdown + r['d']['steps']} of 300 "
                                  f"(t = {r['t']:.4e} s at the end)"},
        "newton_per_stage": r["newton_per_stage"], "krylov_per_newton": r["krylov_per_newton"], "krylov_per_step": r["krylov_per_step"],
        "hbm_step_frac": hbm_frac, "kernel_ms": kernel_ms, "forms": forms_out,
        "roofline": roof,
    }
    if not a.profile:
        log("conservation check")
        out["conservation"] = conservation_check(a, forms)
    if not a.no_cpu and not a.profile:
        if st_plain is None:
            _, st_plain = run_form(make, "plain", a.warmup, 0, 0)
            cfg_plain = st_plain.cfg
        log("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(a, cfg_plain, st_plain)
    if st_plain is not None:
        st_plain.close()
    return out


def _leg(a, make, warmup, steps, workload):
    """one leg: every form over the same window, rates and field errors against the plain FP64 form"""
    res, states = {}, {}
    for form in _forms(a):
        r, st = run_form(make, form, warmup, steps, 0)
        states[form] = st.state()
        st.close()
        res[form] = {k: round(r[k], 4) if isinstance(r[k], float) else r[k]
                     for k in ("value", "ms_per_step", "newton_per_stage", "krylov_per_newton", "krylov_per_step")}
        if "guess" in r:
            res[form]["guess"] = r["guess"]
    for form in res:
        if form != "plain" and "plain" in states:
            res[form]["field_l1_vs_fp64"] = field_l1(states[form], states["plain"])
            res[form]["speedup_vs_plain"] = res[form]["value"] / res["plain"]["value"]
    res["workload"] = workload
    res["window"] = f"steps {warmup + 1}..{warmup + steps}"
    return res


def run_bubble(a):
    """2048x1024 shock-bubble (Mach 1.22 in gas gamma 1.4 onto a gamma-1.648 light-gas bubble, walls top and bottom),
    Newton tolerance 1e-13, timed in a window after the shock has reached the bubble"""
    from diffusionimex_amd import cases
    return _leg(a, lambda **kw: cases.shock_bubble(2048, 1024, seed=a.seed, cfl=0.5, **kw), a.bubble_warmup, a.bubble_steps,
                "2D shock-bubble 2048x1024 on 0.2x0.1 m, Newton tol 1e-13 (the shock reaches the bubble near step 250)")


def run_interface(a):
    """4096x2048 Mach-1.5 shock on a perturbed water-gas interface (seeded amplitude and mode), Newton tolerance 1e-10"""
    from diffusionimex_amd import cases
    return _leg(a, lambda **kw: cases.perturbed_interface(4096, 2048, seed=a.seed, cfl=0.5, newton_tol=1e-10, **kw),
                a.interface_warmup, a.interface_steps,
                "4096x2048 shock on a perturbed water-gas interface, Newton tol 1e-10 (the shock travels in gas during the "
                "500 steps of the workload)")


def run_batch(a):
    """Batched 1D leg: 4096 two-phase Riemann problems (lengths 256..8192, per-problem stiffened-gas liquid, p_L, p_R
    log-uniform in [1e5, 1e9] Pa) advanced by one persistent launch per IMEX stage, replayed as a hipGraph, with the
    autotuned launch shape; cell-updates/s over the timed steps after the warm-up, and the oracle on the host cores over
    a sample of the problems resumed from the same warmed state."""
    from diffusionimex_amd import cases
    from diffusionimex_amd.batch import Batch
    scheme, probs, Us, _ = cases.riemann_batch(4096, seed=a.seed)
    b = Batch(probs, scheme)
    b.set_state(Us)
    if b.run(1.0, a.batch_warmup) != 0:
        raise SystemExit("batch warm-up failed")
    tune = b.autotune(steps=4)
    per0, tot0 = b.stats()
    t0 = time.perf_counter()
    rc = b.run(1.0, a.batch_steps)
    wall = time.perf_counter() - t0
    per1, tot1 = b.stats()
    if rc != 0:
        raise SystemExit("batch timed run failed")
    cells = b.cells
    d = {k: tot1[k] - tot0[k] for k in ("steps", "stages", "newton_it", "rejects")}
    res = {"workload": f"{len(probs)} batched 1D Riemann problems, {cells} cells (lengths 256..8192), Newton tol 1e-10, CFL 0.5",
           "metric": "cell-updates/s", "value": cells * a.batch_steps / wall, "steps": a.batch_steps, "warmup": a.batch_warmup,
           "ms_per_step": 1e3 * wall / a.batch_steps, "newton_per_stage": d["newton_it"] / max(d["stages"], 1),
           "rejects": d["rejects"], "launch": {"threads": tune["threads"], "per_cu": tune["per_cu"]},
           "autotune_ms": {f"{t}x{c}": round(ms, 1) for t, c, ms in tune["tried"]}}
    if not a.no_cpu and not a.profile:
        res["cpu_baseline"] = batch_cpu_baseline(a, scheme, probs, b, per1)
    b.close()
    return res


def _physical_cores():
    """(physical cores among the CPUs this process may run on, one CPU id per core)"""
    try:
        allowed = sorted(os.sched_getaffinity(0))
    except AttributeError:
        allowed = list(range(os.cpu_count() or 1))
    seen, pick = set(), []
    for c in allowed:
        try:
            base = f"/sys/devices/system/cpu/cpu{c}/topology/"
            key = (open(base + "physical_package_id").read().strip(), open(base + "core_id").read().strip())
        except OSError:
            key = ("?", str(c))
        if key not in seen:
            seen.add(key)
            pick.append(c)
    return len(pick), pick


def _pin_cpu_threads(a):
    """one OpenMP thread per physical core (at most 16, this box's CPU share for one GPU), pinned to cores; set before
    the oracle's OpenMP runtime starts"""
    ncore, pick = _physical_cores()
    threads = a.cpu_threads or max(1, min(16, ncore))
    a.cpu_threads_used = threads
    a.cpu_cores_available = ncore
    os.environ["OMP_PLACES"] = "{" + "},{".join(str(c) for c in pick[:threads]) + "}"
    os.environ["OMP_PROC_BIND"] = "close"


def batch_cpu_baseline(a, scheme, probs, b, per):
    """The oracle (PC_omp paths) on the host cores, one problem per thread: a sample of the batch's problems resumed from
    the GPU's warmed state, batch-steps steps each."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    from tests.helpers import oracle_cfg_from
    from diffusionimex_amd import spb6
    threads = a.cpu_threads_used
    O.lib().ora_set_threads(1)
    U = b.state()
    P, pref = b.pressure()
    pick = np.linspace(0, len(probs) - 1, min(a.batch_cpu_problems, len(probs))).astype(int)

    def one(p):
        pr = probs[p]
        ph = [tuple(getattr(pr.ph[k], f) for f in ("gamma", "pi_inf", "q", "qp", "cv", "b")) for k in range(2)]
        cfg = spb6.make_config(pr.nx, 1, phases=ph, domain=(pr.x0, pr.x1, 0.0, 1.0), cfl=scheme.cfl, newton_tol=scheme.newton_tol)
        oc = oracle_cfg_from(cfg, O)
        O.reference_mode(oc, "PC_omp")
        o = O.Oracle(oc)
        o.set_state(b.problem_state(U, p).reshape(1, -1, 8))
        o.set_pressure(P[b.off[p]:b.off[p + 1]].reshape(1, -1), pref[p])
        s = per[p]
        o.set_clock(s["t"], s["dt_last"], 1 if s["rk1"] >= s["steps"] else 0)
        rc = o.run(1.0, a.batch_steps)
        return pr.nx * o.stats()["steps"], rc

    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        outs = list(ex.map(one, pick))
    wall = time.perf_counter() - t0
    cu = sum(c for c, _ in outs)
    return {"value": cu / wall, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{len(pick)} of the {len(probs)} problems, {a.batch_steps} steps each from the GPU's warmed state, "
                      f"one problem per host thread ({wall:.1f} s, {sum(1 for _, r in outs if r != 0)} failed)"}


def cpu_baseline(a, cfg, st):
    """The oracle (PC_omp paths, B = 1024 block sums, libm) on the host's physical cores, threads pinned: cpu-steps
    LSDIRK2 steps of the same grid resumed from the GPU's warmed plain-form state (same t, dt history, pressure)."""
    from oracle import oracle as O
    from tests.helpers import oracle_cfg_from
    threads = a.cpu_threads_used
    L = O.lib()
    L.ora_set_threads(threads)
    oc = oracle_cfg_from(cfg, O)
    O.reference_mode(oc, "PC_omp")
    o = O.Oracle(oc)
    o.set_state(st.state())
    P, pref = st.pressure()
    o.set_pressure(P, pref)
    s = st.stats()
    o.set_clock(st.t, s["dt_last"], 0)
    t0 = time.perf_counter()
    rc = o.run(1.0, a.cpu_steps)
    t1 = time.perf_counter()
    n = cfg.nx * cfg.ny
    so = o.stats()
    return {"value": n * so["steps"] / (t1 - t0), "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{so['steps']} LSDIRK2 step(s) of the {cfg.nx}x{cfg.ny} workload resumed from the GPU plain-form state "
                      f"after {s['steps']} steps ({t1 - t0:.1f} s, rc {rc}, {so['newton_it']} Newton, {so['krylov_it']} BiCGStab)",
            "cpu": _cpu_model(), "physical_cores_available": a.cpu_cores_available, "pinned": os.environ.get("OMP_PROC_BIND") == "close"}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
