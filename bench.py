"""Benchmark: 2D IMEX cell-updates/s of the SPB 6+1 semi-implicit time step on one MI355X.

Workload: 2D shock-bubble, 2048 x 1024 cells on 0.2 x 0.1 m (Mach-1.22 shock in gas gamma 1.4 onto a gamma-1.648
light-gas bubble, reflective top/bottom walls), FP64, material CFL 0.5, LSDIRK2, Newton tol 1e-13, BiCGStab tol 1e-13.
A "step" is one accepted LSDIRK2 time step (2 stages with all their Newton and BiCGStab work). --warmup steps run first
and cover the start-up ramp (the first 17 steps of this workload are first-order steps whose dt grows from the acoustic
step to the material CFL step), then exactly --steps full LSDIRK2 steps are timed; every stepper call returns after its HIP stream
synchronises, so the timed region is bracketed by device synchronisation on both sides.

Prints one JSON line with the metric, the roofline of the dominant kernel (algorithmic bytes per launch / average launch
duration from HIP events on the stepper's stream) and the CPU baseline (the oracle's PC_omp-path restatement on the host
cores, a bounded sample of the same workload resumed from the GPU's warmed state).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12   # MI355X HBM3E, B/s
# algorithmic bytes per interior cell per launch (FP64, 2D): compulsory reads + writes of each kernel's data flow
BYTES_2D = {"explicit": 232, "prep": 184, "residual": 256, "jacobian": 304, "krylov_iter": 296, "final": 376}
# one BiCGStab iteration of the FP32 form (red-black sweep preconditioner, five FP32 passes), bytes per row: kb_p 40
# (r, p, v, 1/d and the four off-diagonals in; new p and M^-1 p out), kb_v 32 (five diagonals, M^-1 p, r0 in; v out),
# kb_s 36 (r, v, 1/d and the four off-diagonals in; s, M^-1 s out), kb_t 32 (five diagonals, M^-1 s, s in; t out),
# kb_xr 32 (M^-1 p, M^-1 s, s, t, x, r0 in; x, r out); a black row's red neighbours come through the caches
BYTES_KRYLOV_FP32 = 172


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=2048)
    ap.add_argument("--ny", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile", action="store_true", help="short run for rocprofv3 (no CPU legs; legs default to the 2D one)")
    ap.add_argument("--legs", default=None,
                    help="comma list of shock-bubble (the headline 2048x1024 leg), interface, batch; default all three "
                         "(shock-bubble only with --profile)")
    ap.add_argument("--interface-steps", type=int, default=6,
                    help="timed steps of the 4096x2048 perturbed-interface leg, plain vs learned + FP32 Krylov (0: skip)")
    ap.add_argument("--interface-warmup", type=int, default=20,
                    help="warm-up steps of the interface leg (its start-up ramp is 15 first-order steps)")
    ap.add_argument("--batch-steps", type=int, default=8, help="timed steps of the batched 1D leg (0: skip the leg)")
    ap.add_argument("--batch-warmup", type=int, default=12)
    ap.add_argument("--batch-cpu-problems", type=int, default=64)
    ap.add_argument("--forms", default=None,
                    help="comma list of plain, mixed (FP32 Krylov), learned (MLP guess + FP32 Krylov); default all three "
                         "when the trained network is present, else plain,mixed")
    return ap.parse_args()


def run_form(a, form, keep_state=False):
    """warm up and time one form of the stepper on the workload; returns (result dict, stepper, cfg)"""
    from diffusionimex_amd import spb6, cases
    kw = {}
    if form in ("mixed", "learned"):
        kw["krylov_fp32"] = 1
        kw["analytic_jacobian"] = 1
    if form == "learned":
        kw["learned_guess"] = 1
    cfg, U = cases.shock_bubble(a.nx, a.ny, seed=a.seed, cfl=0.5, **kw)
    n = a.nx * a.ny
    st = spb6.Stepper(cfg, device=0)
    if form == "learned":
        from diffusionimex_amd import mlp
        mlp.load_into(st)
    st.set_state(U)
    rc = st.run(1.0, a.warmup)
    if rc != 0:
        raise SystemExit(f"{form}: warmup failed: {spb6.STATUS.get(rc, rc)}")
    keys = ("steps", "stages", "newton_it", "krylov_it", "linsolves", "guess_used", "guess_fallback")
    # the per-category breakdown: K steps with every kernel category bracketed by HIP events (some 180 event records a
    # step, which cost about 8% of the step), ahead of the timed region
    s0 = st.stats()
    st.enable_timing(True)
    rc = st.run(1.0, a.steps)
    if rc != 0:
        raise SystemExit(f"{form}: breakdown run failed: {spb6.STATUS.get(rc, rc)}")
    kt_all = st.kernel_times()
    s1 = st.stats()
    d_all = {k: s1[k] - s0[k] for k in keys}
    dom = dominant(kt_all)
    # the timed region: K more steps, only the dominant category's launches bracketed by events (its live duration)
    st.enable_timing(True, categories=TIMER_CATS[dom])
    t0 = time.perf_counter()
    rc = st.run(1.0, a.steps)
    t1 = time.perf_counter()
    if rc != 0:
        raise SystemExit(f"{form}: timed run failed: {spb6.STATUS.get(rc, rc)}")
    s2 = st.stats()
    kt_live = st.kernel_times()
    st.enable_timing(False)
    wall = t1 - t0
    d = {k: s2[k] - s1[k] for k in keys}
    res = {"value": n * d["steps"] / wall, "wall": wall, "ms_per_step": 1e3 * wall / max(d["steps"], 1),
           "newton_per_stage": d["newton_it"] / max(d["stages"], 1), "krylov_per_newton": d["krylov_it"] / max(d["linsolves"], 1),
           "krylov_per_step": d["krylov_it"] / max(d["steps"], 1), "d": d, "kt": kt_all, "d_all": d_all, "dom": dom, "kt_live": kt_live}
    if form == "learned":
        res["guess_used"], res["guess_fallback"] = d["guess_used"], d["guess_fallback"]
    return res, st, cfg


# HBM bytes per launch of the head form's candidate dominant kernels at 2048x1024, measured by rocprofv3 --pmc FETCH_SIZE
# and --pmc WRITE_SIZE, calibrated on the FP32 conversion kernel (known bytes: FETCH_SIZE reads x 2.00, WRITE_SIZE
# x 1.00): one FP32 BiCGStab iteration (the five passes' counts over the run divided by the iterations taken) and one
# residual evaluation writing its trial pressure (median active dispatch); commands and per-pass counts in
# profiles/krylov_fp32_iteration_2048x1024_traffic.md and profiles/residual_2048x1024_traffic.md
TRAFFIC = {("krylov_iter", True, 2048, 1024): 327036314, ("residual", True, 2048, 1024): 782086963}


# timer categories (Stepper.TIMING_CATEGORIES) behind each roofline category
TIMER_CATS = {"explicit": ("explicit",), "prep": ("prep",), "residual": ("residual",), "jacobian": ("jacobian",),
              "krylov_iter": ("krylov", "krylov_idle"), "final": ("final",)}


def categories(kt, krylov_it, n, fp32):
    """roofline category -> (timed ms, launches, algorithmic bytes per launch)"""
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
    """the category with the most kernel time"""
    cats = categories(kt, 1, 1, False)
    return max(cats, key=lambda k: cats[k][0])


def roofline_of(res, n, fp32, nx=2048, ny=1024):
    kt, d = res["kt"], res["d"]
    kb = BYTES_KRYLOV_FP32 if fp32 else BYTES_2D["krylov_iter"]
    cats = categories(kt, res["d_all"]["krylov_it"], n, fp32)
    kernel_ms = {k: round(v[0], 3) for k, v in cats.items()}
    kernel_ms["shock"] = round(kt["shock_ms"], 3)
    kernel_ms["other"] = round(kt["other_ms"], 3)
    kernel_ms["krylov_aux"] = round(kt["krylov_aux_ms"], 3)
    kernel_ms["krylov_idle"] = round(kt["krylov_idle_ms"], 3)
    kernel_ms["guess"] = round(kt["guess_ms"], 3)
    # the dominant category's launches as timed live in the timed region
    dom = res["dom"]
    ms, cnt, byts = categories(res["kt_live"], d["krylov_it"], n, fp32)[dom]
    avg_s = (ms / 1e3) / max(cnt, 1)
    achieved = byts / avg_s if avg_s > 0 else 0.0
    traffic = TRAFFIC.get((dom, fp32, nx, ny))
    roof = {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": traffic, "bytes_per_launch": byts, "avg_launch_ms": avg_s * 1e3,
            "iterations": cnt,
            # every category on its algorithmic bytes over its timed region (the region's small boundary-fill and
            # reduction kernels included)
            "by_kernel": {k: round(v[2] * max(v[1], 1) / (v[0] / 1e3) / HBM_PEAK, 4) for k, v in cats.items() if v[0] > 0}}
    step_bytes = (BYTES_2D["explicit"] + BYTES_2D["prep"] + BYTES_2D["final"]) * d["stages"] + BYTES_2D["jacobian"] * d["newton_it"] \
        + BYTES_2D["residual"] * (d["stages"] + d["newton_it"]) + kb * d["krylov_it"]
    return roof, kernel_ms, step_bytes * n / res["wall"] / HBM_PEAK


def main():
    a = parse()
    if a.legs is None:
        a.legs = "shock-bubble" if a.profile else "shock-bubble,interface,batch"
    legs = set(x for x in a.legs.split(",") if x)
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}); multi-GPU sharding is design-only", file=sys.stderr)
        sys.exit(2)
    out = run_headline(a) if "shock-bubble" in legs else {"legs": sorted(legs)}
    if a.interface_steps > 0 and "interface" in legs:
        out["interface"] = run_interface(a)
    if a.batch_steps > 0 and "batch" in legs:
        out["batch"] = run_batch(a)
    print(json.dumps(out))


def run_headline(a):
    """The 2048x1024 shock-bubble leg in each form (plain FP64, FP32 Krylov, learned guess + FP32 Krylov): the JSON
    contract fields, with the head form's dominant-kernel roofline and the oracle CPU baseline."""
    n = a.nx * a.ny
    if a.forms is None:
        from diffusionimex_amd import mlp
        a.forms = "plain,mixed,learned" if os.path.exists(mlp.WEIGHTS) else "plain,mixed"
    forms = [f for f in a.forms.split(",") if f]
    results, states = {}, {}
    st_plain = cfg = None
    for form in forms:
        res, st, c = run_form(a, form)
        states[form] = st.state()
        results[form] = res
        if form == "plain":
            st_plain, cfg = st, c
        else:
            st.close()
        if cfg is None:
            cfg = c
    head = "learned" if "learned" in results else ("mixed" if "mixed" in results and "plain" not in results else "plain")
    r = results[head]
    roof, kernel_ms, hbm_frac = roofline_of(r, n, head != "plain", a.nx, a.ny)
    forms_out = {}
    for form, res in results.items():
        fo = {k: res[k] for k in ("value", "ms_per_step", "newton_per_stage", "krylov_per_newton", "krylov_per_step")}
        if "plain" in states and form != "plain":
            A, B = states[form].reshape(-1, 8), states["plain"].reshape(-1, 8)
            fo["field_l1_vs_fp64"] = [float(np.abs(A[:, k] - B[:, k]).sum() / max(np.abs(B[:, k]).sum(), 1e-300)) for k in range(8)]
            fo["speedup_vs_plain"] = res["value"] / results["plain"]["value"]
        if form == "learned":
            fo["guess_used"], fo["guess_fallback"] = res["guess_used"], res["guess_fallback"]
        forms_out[form] = fo
    out = {
        "metric": "2D IMEX cell-updates/s + BiCGStab iters/step, learned+mixed vs plain FP64",
        "value": r["value"], "unit": "cell-updates/s", "n_gpus": 1, "steps": r["d"]["steps"], "warmup": a.warmup,
        "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if head == "plain" else "f64 (FP32 Krylov)", "data": "synthetic",
        "config": {"workload": f"2D shock-bubble {a.nx}x{a.ny} ({head} form; Mach 1.22, gamma 1.4 / 1.648, walls top/bottom)",
                   "cells": n, "seed": a.seed, "cfl": 0.5, "newton_tol": 1e-13, "krylov_tol": 1e-13, "form": head},
        "newton_per_stage": r["newton_per_stage"], "krylov_per_newton": r["krylov_per_newton"], "krylov_per_step": r["krylov_per_step"],
        "hbm_step_frac": hbm_frac, "kernel_ms": kernel_ms, "forms": forms_out,
        "roofline": roof,
    }
    if not a.no_cpu and not a.profile:
        if st_plain is None:
            _, st_plain, cfg = run_form(a, "plain")
        out["cpu_baseline"] = cpu_baseline(a, cfg, st_plain)
    if st_plain is not None:
        st_plain.close()
    return out


def run_interface(a):
    """4096x2048 Mach-1.5 shock on a perturbed water-gas interface (seeded amplitude and mode), Newton tolerance 1e-10:
    the plain FP64 form against the learned guess + FP32 Krylov form, same warm-up and timed steps."""
    from diffusionimex_amd import spb6, cases, mlp
    res, states = {}, {}
    forms = ["plain", "learned"] if os.path.exists(mlp.WEIGHTS) else ["plain", "mixed"]
    for form in forms:
        kw = {} if form == "plain" else {"krylov_fp32": 1, "analytic_jacobian": 1}
        if form == "learned":
            kw["learned_guess"] = 1
        cfg, U = cases.perturbed_interface(4096, 2048, seed=a.seed, cfl=0.5, newton_tol=1e-10, **kw)
        st = spb6.Stepper(cfg, device=0)
        if form == "learned":
            mlp.load_into(st)
        st.set_state(U.reshape(cfg.ny, cfg.nx, 8))
        if st.run(1.0, a.interface_warmup) != 0:
            raise SystemExit(f"interface {form}: warm-up failed")
        s0 = st.stats()
        t0 = time.perf_counter()
        rc = st.run(1.0, a.interface_steps)
        wall = time.perf_counter() - t0
        s1 = st.stats()
        if rc != 0:
            raise SystemExit(f"interface {form}: timed run failed")
        d = {k: s1[k] - s0[k] for k in ("steps", "stages", "newton_it", "krylov_it", "linsolves")}
        res[form] = {"value": cfg.nx * cfg.ny * d["steps"] / wall, "ms_per_step": 1e3 * wall / d["steps"],
                     "newton_per_stage": d["newton_it"] / max(d["stages"], 1),
                     "krylov_per_newton": d["krylov_it"] / max(d["linsolves"], 1),
                     "krylov_per_step": d["krylov_it"] / max(d["steps"], 1)}
        states[form] = st.state().reshape(-1, 8)
        st.close()
    other = forms[1]
    A, B = states[other], states["plain"]
    res[other]["field_l1_vs_plain"] = [float(np.abs(A[:, k] - B[:, k]).sum() / max(np.abs(B[:, k]).sum(), 1e-300)) for k in range(8)]
    res[other]["speedup_vs_plain"] = res[other]["value"] / res["plain"]["value"]
    res["workload"] = f"4096x2048 shock on a perturbed water-gas interface, Newton tol 1e-10, {a.interface_warmup} warm-up steps"
    res["steps"] = a.interface_steps
    return res


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


def batch_cpu_baseline(a, scheme, probs, b, per):
    """The oracle (PC_omp paths) on the host cores, one problem per thread: a sample of the batch's problems resumed from
    the GPU's warmed state, batch-steps steps each."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    from tests.helpers import oracle_cfg_from
    from diffusionimex_amd import spb6
    threads = a.cpu_threads or min(16, os.cpu_count() or 1)
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
    """The oracle (PC_omp paths, B = 1024 block sums, libm) on the host cores: cpu-steps LSDIRK2 steps of the same grid
    resumed from the GPU's warmed state (same t, dt history, pressure)."""
    from oracle import oracle as O
    from tests.helpers import oracle_cfg_from
    threads = a.cpu_threads or min(16, os.cpu_count() or 1)
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
            "sample": f"{so['steps']} LSDIRK2 steps of the {cfg.nx}x{cfg.ny} workload resumed from the GPU state after "
                      f"{a.warmup + 2 * a.steps} steps ({t1 - t0:.1f} s, rc {rc}, {so['newton_it']} Newton, {so['krylov_it']} BiCGStab)",
            "cpu": _cpu_model()}


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
