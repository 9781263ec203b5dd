"""Python binding of the gfx950 SPB 6+1 IMEX stepper (libspb6_hip.so, C ABI in include/spb6_imex.h).

The class Stepper mirrors the reference solver's stepper interface (src/SPB6_PC_IMEX.cpp): the state grid and pressure
field, caldt, bound_cond_u, TVDReconstruction + Update (explicit stage), PressureSubsystem (implicit stage), one step
of the attempt loop and the time loop. Every call runs on the GPU; there is no CPU fallback -- loading fails loudly
when the HIP library is missing.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspb6_hip.so")
LIB_MG_PATH = os.path.join(HERE, "libspb6_hip_mg.so")   # the same library built with the Mie-Grueneisen EOS branches

BC_TRANSMISSIVE, BC_WALL, BC_PERIODIC = 0, 1, 2
STATUS = {0: "OK", 1: "RHO_NONPOSITIVE", 2: "STAGE_INADMISSIBLE", 3: "CLOSURE_FAIL", 4: "NEWTON_FAIL",
          5: "KRYLOV_FAIL", 6: "DT_UNDERFLOW", 7: "STEP_FAILED", 8: "BAD_ARG", 9: "HIP_ERROR"}
STATUS_CODE = {v: k for k, v in STATUS.items()}
# IMEX variants (spb6_variant): the reference's sibling programs SPB6_<name>_IMEX.cpp
VARIANTS = {"PC": 0, "TV": 1, "TVpure": 2, "TVinc": 3, "TVsplit": 4, "TVbb": 5, "BB": 6}


class Phase(C.Structure):
    _fields_ = [("gamma", C.c_double), ("pi_inf", C.c_double), ("q", C.c_double), ("qp", C.c_double),
                ("cv", C.c_double), ("b", C.c_double)]


class Config(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int),
        ("x0", C.c_double), ("x1", C.c_double), ("y0", C.c_double), ("y1", C.c_double),
        ("bc", C.c_int * 4), ("bcu", C.c_double * 4), ("ph", Phase * 2),
        ("rk", C.c_int), ("cfl", C.c_double), ("cfl_ac", C.c_double), ("ramp", C.c_int), ("dtfix", C.c_double),
        ("acdt", C.c_int), ("newton_tol", C.c_double), ("newton_iter", C.c_int),
        ("krylov_tol", C.c_double), ("krylov_iter", C.c_int),
        ("limiter", C.c_int), ("order", C.c_int), ("rus_local", C.c_int),
        ("krylov_fp32", C.c_int), ("learned_guess", C.c_int), ("analytic_jacobian", C.c_int),
        ("jreuse", C.c_int), ("variant", C.c_int), ("relax", C.c_int), ("eps_i", C.c_double), ("graph", C.c_int),
        ("shock_fused", C.c_int), ("theta_rate", C.c_double), ("nu_rate", C.c_double),
        ("eos_kind", C.c_int * 2), ("mg", (C.c_double * 7) * 2), ("mu_rate", C.c_double),
        ("nu_law", C.c_int), ("nu_cr", C.c_double), ("nu_pc", C.c_double), ("relax_form", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [("steps", C.c_long), ("rk1", C.c_long), ("rejects", C.c_long), ("floor_acc", C.c_long),
                ("stages", C.c_long), ("newton_it", C.c_long), ("krylov_it", C.c_long), ("linsolves", C.c_long),
                ("maxres", C.c_double), ("t", C.c_double), ("dt_last", C.c_double), ("status", C.c_int),
                ("guess_used", C.c_long), ("guess_fallback", C.c_long), ("guess_log10_gain", C.c_double),
                ("guess_skipped", C.c_long), ("tmax", C.c_double * 2), ("closure_fails", C.c_long),
                ("jac_builds", C.c_long)]


class KernelTimes(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("explicit_ms", "prep_ms", "shock_ms", "residual_ms", "jacobian_ms",
                                           "krylov_ms", "final_ms", "other_ms")] + \
               [(k, C.c_long) for k in ("explicit_n", "prep_n", "shock_n", "residual_n", "jacobian_n", "krylov_n",
                                         "final_n", "krylov_iters")] + \
               [("krylov_aux_ms", C.c_double), ("guess_ms", C.c_double), ("krylov_launched", C.c_long),
                ("krylov_idle_ms", C.c_double), ("shock_rk_steps", C.c_long), ("graph_launches", C.c_long),
                ("graph_rebuilds", C.c_long)]


# every symbol include/spb6_imex.h declares
EXPORTS = ["spb6_create", "spb6_destroy", "spb6_last_error", "spb6_set_state", "spb6_get_state", "spb6_get_pressure",
           "spb6_set_pressure", "spb6_caldt", "spb6_bc_fill", "spb6_explicit_stage", "spb6_implicit_stage",
           "spb6_stage_prep", "spb6_step", "spb6_attempt_begin", "spb6_stage_combine", "spb6_stage_update",
           "spb6_stage_pressure", "spb6_attempt_accept", "spb6_run", "spb6_relax_activate", "spb6_hllc_step", "spb6_hllc_run",
           "spb6_get_profile", "spb6_write_profile",
           "spb6_conserved_totals", "spb6_stats_line", "spb6_get_stats", "spb6_time", "spb6_residual",
           "spb6_jacobian", "spb6_linsolve", "spb6_linsolve_device", "spb6_get_field", "spb6_enable_timing", "spb6_set_timing_mask", "spb6_get_kernel_times",
           "spb6_mlp_create", "spb6_mlp_destroy", "spb6_mlp_nparams", "spb6_mlp_set_params", "spb6_mlp_get_params",
           "spb6_mlp_set_data", "spb6_mlp_train", "spb6_mlp_forward", "spb6_mlp_gradients", "spb6_mlp_infer_bf16",
           "spb6_set_guess_model", "spb6_guess_residual", "spb6_dump_enable", "spb6_dump_fetch",
           "spb6_batch_create", "spb6_batch_destroy", "spb6_batch_last_error", "spb6_batch_cells", "spb6_batch_set_state",
           "spb6_batch_get_state", "spb6_batch_get_pressure", "spb6_batch_run", "spb6_batch_get_stats",
           "spb6_batch_set_launch", "spb6_batch_autotune"]

_libs = {}


def load(mg=False):
    """Load libspb6_hip.so, or with mg its Mie-Grueneisen build libspb6_hip_mg.so (both built by
    diffusionimex_amd/build.py); raises if it is missing."""
    path = LIB_MG_PATH if mg else LIB_PATH
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library {path} is missing: run `python diffusionimex_amd/build.py` (no CPU fallback)")
    L = C.CDLL(path)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.spb6_create.restype = vp
    L.spb6_create.argtypes = [C.POINTER(Config), C.c_int]
    L.spb6_destroy.argtypes = [vp]
    L.spb6_last_error.restype = C.c_char_p
    L.spb6_set_state.argtypes = [vp, dp]
    L.spb6_get_state.argtypes = [vp, dp]
    L.spb6_get_pressure.argtypes = [vp, dp, dp]
    L.spb6_set_pressure.argtypes = [vp, dp, C.c_double]
    L.spb6_caldt.argtypes = [vp, C.c_int, dp]
    L.spb6_bc_fill.argtypes = [vp]
    L.spb6_explicit_stage.argtypes = [vp, C.c_double]
    L.spb6_implicit_stage.argtypes = [vp, C.c_double, ip, dp]
    L.spb6_stage_prep.argtypes = [vp, C.c_double]
    L.spb6_step.argtypes = [vp, C.c_double, C.c_int]
    L.spb6_attempt_begin.argtypes = [vp]
    L.spb6_stage_combine.argtypes = [vp, C.c_int, C.c_double, dp, dp]
    L.spb6_stage_update.argtypes = [vp, C.c_double]
    L.spb6_stage_pressure.argtypes = [vp, C.c_double, C.c_int, ip, dp]
    L.spb6_attempt_accept.argtypes = [vp, C.c_double]
    L.spb6_get_profile.argtypes = [vp, dp]
    L.spb6_write_profile.argtypes = [vp, C.c_char_p, C.c_int]
    L.spb6_conserved_totals.argtypes = [vp, dp]
    L.spb6_stats_line.argtypes = [vp, C.c_int, C.c_double, C.c_char_p, C.c_long]
    L.spb6_run.argtypes = [vp, C.c_double, C.c_long]
    L.spb6_hllc_step.argtypes = [vp, C.c_double]
    L.spb6_relax_activate.argtypes = [vp]
    L.spb6_hllc_run.argtypes = [vp, C.c_double, C.c_long, C.c_double]
    L.spb6_get_stats.argtypes = [vp, C.POINTER(Stats)]
    L.spb6_time.restype = C.c_double
    L.spb6_time.argtypes = [vp]
    L.spb6_residual.argtypes = [vp, dp, dp]
    L.spb6_jacobian.argtypes = [vp, dp, dp, dp, dp]
    L.spb6_linsolve.argtypes = [vp, dp, dp, dp, C.c_int, ip]
    L.spb6_linsolve_device.argtypes = [C.c_int, C.c_int, ip, vp, vp, vp, C.c_double, C.c_int, C.c_int, ip, vp]
    L.spb6_get_field.restype = C.c_long
    L.spb6_get_field.argtypes = [vp, C.c_char_p, dp, C.c_long]
    L.spb6_enable_timing.argtypes = [vp, C.c_int]
    L.spb6_set_timing_mask.argtypes = [vp, C.c_uint]
    L.spb6_get_kernel_times.argtypes = [vp, C.POINTER(KernelTimes)]
    fp = C.POINTER(C.c_float)
    L.spb6_mlp_create.restype = vp
    L.spb6_mlp_create.argtypes = [C.c_int]
    L.spb6_mlp_destroy.argtypes = [vp]
    L.spb6_mlp_set_params.argtypes = [vp, fp, fp]
    L.spb6_mlp_get_params.argtypes = [vp, fp]
    L.spb6_mlp_set_data.argtypes = [vp, fp, fp, C.c_long]
    L.spb6_mlp_train.argtypes = [vp, C.c_int, C.c_int, C.c_float, C.c_ulonglong, fp]
    L.spb6_mlp_forward.argtypes = [vp, fp, C.c_int, fp]
    L.spb6_mlp_gradients.argtypes = [vp, fp, fp, C.c_int, fp, fp]
    L.spb6_mlp_infer_bf16.argtypes = [vp, fp, C.c_int, fp]
    L.spb6_set_guess_model.argtypes = [vp, vp]
    L.spb6_guess_residual.argtypes = [vp, dp, dp, dp, dp, dp, dp]
    L.spb6_dump_enable.argtypes = [vp, C.c_long, C.c_int]
    L.spb6_dump_fetch.restype = C.c_long
    L.spb6_dump_fetch.argtypes = [vp, fp, fp, C.c_long]
    L.spb6_batch_create.restype = vp
    L.spb6_batch_create.argtypes = [vp, C.c_int, C.POINTER(Config), C.c_int]
    L.spb6_batch_destroy.argtypes = [vp]
    L.spb6_batch_last_error.restype = C.c_char_p
    L.spb6_batch_last_error.argtypes = [vp]
    L.spb6_batch_cells.restype = C.c_long
    L.spb6_batch_cells.argtypes = [vp]
    L.spb6_batch_set_state.argtypes = [vp, dp]
    L.spb6_batch_get_state.argtypes = [vp, dp]
    L.spb6_batch_get_pressure.argtypes = [vp, dp, dp]
    L.spb6_batch_run.argtypes = [vp, C.c_double, C.c_long]
    L.spb6_batch_get_stats.argtypes = [vp, C.POINTER(Stats), C.POINTER(Stats)]
    L.spb6_batch_set_launch.argtypes = [vp, C.c_int, C.c_int]
    L.spb6_batch_autotune.argtypes = [vp, C.c_long, vp]
    _libs[path] = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_config(nx, ny=1, phases=((4.4, 6e8), (1.4, 0.0)), domain=(0.0, 1.0, 0.0, 1.0), bc=(0, 0, 0, 0), bcu=(0, 0, 0, 0),
                rk=2, cfl=0.2, cfl_ac=0.4, ramp=1, dtfix=0.0, acdt=0, newton_tol=1e-13, newton_iter=60,
                krylov_tol=1e-13, krylov_iter=4000, limiter=1, order=2, rus_local=1, krylov_fp32=0, learned_guess=0,
                analytic_jacobian=0, jreuse=0, variant=0, relax=0, eps_i=0.0, graph=0, shock_fused=0, theta_rate=0.0, nu_rate=0.0,
                eos_kind=(0, 0), mg=((0.0,) * 7, (0.0,) * 7), mu_rate=0.0,
                nu_law=0, nu_cr=0.0, nu_pc=0.0, relax_form=0):
    """Config with the reference defaults (PC:57-80): phases are (gamma, p_inf[, q, qp, cv, b]); eos_kind[k] 1 or 2 makes
    phase k a Cochran-Chan or JWL phase with mg[k] = (G, c1, c2, e1, e2, rr, e0)."""
    c = Config()
    c.nx, c.ny = nx, ny
    c.x0, c.x1, c.y0, c.y1 = domain
    for q in range(4):
        c.bc[q] = bc[q]
        c.bcu[q] = bcu[q]
    for k, p in enumerate(phases):
        full = list(p) + [0.0] * (6 - len(p))
        if len(p) < 5:
            full[4] = 1000.0
        c.ph[k] = Phase(*full)
    c.rk, c.cfl, c.cfl_ac, c.ramp, c.dtfix, c.acdt = rk, cfl, cfl_ac, ramp, dtfix, acdt
    c.newton_tol, c.newton_iter, c.krylov_tol, c.krylov_iter = newton_tol, newton_iter, krylov_tol, krylov_iter
    c.limiter, c.order, c.rus_local, c.krylov_fp32, c.learned_guess = limiter, order, rus_local, krylov_fp32, learned_guess
    c.analytic_jacobian = analytic_jacobian
    c.jreuse = jreuse
    c.variant = variant
    c.relax, c.eps_i = relax, eps_i
    c.graph = graph
    c.shock_fused = shock_fused
    c.theta_rate, c.nu_rate = theta_rate, nu_rate
    c.mu_rate = mu_rate
    c.nu_law, c.nu_cr, c.nu_pc = nu_law, nu_cr, nu_pc
    c.relax_form = relax_form
    for k in range(2):
        c.eos_kind[k] = eos_kind[k]
        for q in range(7):
            c.mg[k][q] = mg[k][q]
    return c


class SPB6Error(RuntimeError):
    pass


class Stepper:
    """GPU stepper context: owns all device buffers for one grid (PC:1157-1163, StageData PC:565-582)."""

    def __init__(self, cfg, device=0):
        self.L = load(mg=bool(cfg.eos_kind[0] or cfg.eos_kind[1]))
        self.cfg = cfg
        self.h = self.L.spb6_create(C.byref(cfg), int(device))
        if not self.h:
            raise SPB6Error("spb6_create failed: " + self.L.spb6_last_error().decode())
        self.nx, self.ny = cfg.nx, cfg.ny
        self.NXT, self.NYT = cfg.nx + 4, cfg.ny + 4

    def close(self):
        if getattr(self, "h", None):
            self.L.spb6_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise SPB6Error(f"{what}: {STATUS.get(rc, rc)} {self.L.spb6_last_error().decode()}")
        return rc

    # --- state (u_initialization / bound_cond_u, PC:450-475; pressure init PC:1222-1227)
    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.ny, self.nx, 8)
        self._check(self.L.spb6_set_state(self.h, _dp(U)), "set_state")

    def state(self):
        U = np.empty((self.ny, self.nx, 8))
        self._check(self.L.spb6_get_state(self.h, _dp(U)), "get_state")
        return U

    def pressure(self):
        P = np.empty(self.nx * self.ny)
        pref = np.zeros(1)
        self._check(self.L.spb6_get_pressure(self.h, _dp(P), _dp(pref)), "get_pressure")
        return P.reshape(self.ny, self.nx), float(pref[0])

    def set_pressure(self, P, pref):
        P = np.ascontiguousarray(P, dtype=np.float64).ravel()
        self._check(self.L.spb6_set_pressure(self.h, _dp(P), float(pref)), "set_pressure")

    # --- reference stepper functions
    def caldt(self, acoustic=False):
        """caldt (PC:477): 1 / (max(|u|+c)/dx + max(|v|+c)/dy)."""
        r = np.zeros(1)
        self._check(self.L.spb6_caldt(self.h, int(acoustic), _dp(r)), "caldt")
        return float(r[0])

    def bound_cond_u(self):
        self._check(self.L.spb6_bc_fill(self.h), "bc_fill")

    def explicit_stage(self, adt):
        """TVDReconstruction (PC:349) + Update (PC:251) with Qe = Qs = U; returns the status code."""
        return self.L.spb6_explicit_stage(self.h, float(adt))

    def stage_prep(self, adt):
        return self.L.spb6_stage_prep(self.h, float(adt))

    def implicit_stage(self, adt):
        """PressureSubsystem (PC:851): (status, newton iterations, residual)."""
        it = C.c_int(0)
        res = np.zeros(1)
        rc = self.L.spb6_implicit_stage(self.h, float(adt), C.byref(it), _dp(res))
        return rc, it.value, float(res[0])

    def step(self, dt, rk=2):
        return self.L.spb6_step(self.h, float(dt), int(rk))

    def run(self, t_end=1e300, max_steps=-1):
        return self.L.spb6_run(self.h, float(t_end), int(max_steps))

    def relax_activate(self):
        """RelaxationActivation on the state (PC:537-560); the reference applies it before its output (PC:1300)."""
        return self.L.spb6_relax_activate(self.h)

    def hllc_step(self, dt):
        """One step of the reference's explicit HLLC scheme (src/SPB6_explicit.cpp) of size dt."""
        return self.L.spb6_hllc_step(self.h, float(dt))

    def hllc_run(self, t_end=1e300, max_steps=-1, cfl=0.4):
        """The explicit HLLC scheme's time loop at acoustic CFL cfl (EX:537-566)."""
        return self.L.spb6_hllc_run(self.h, float(t_end), int(max_steps), float(cfl))

    def stats(self):
        s = Stats()
        self._check(self.L.spb6_get_stats(self.h, C.byref(s)), "get_stats")
        out = {k: getattr(s, k) for k, _ in Stats._fields_}
        out["tmax"] = (s.tmax[0], s.tmax[1])
        return out

    # --- output (write_profile PC:1002-1013, conserved_totals PC:1104-1107, stats line PC:1337-1339)
    PROFILE_COLUMNS = ("x", "y", "p", "u", "v", "rho", "alpha2", "Y2", "rho1", "rho2", "c_frozen", "c_wood", "T1", "T2", "p1", "p2",
                       "g1-g2")

    def profile(self):
        out = np.empty((self.ny * self.nx, 17))
        self._check(self.L.spb6_get_profile(self.h, _dp(out)), "get_profile")
        return out

    def write_profile(self, path, prec=16):
        self._check(self.L.spb6_write_profile(self.h, str(path).encode(), int(prec)), "write_profile")

    def conserved_totals(self):
        M = np.zeros(3)
        self._check(self.L.spb6_conserved_totals(self.h, _dp(M)), "conserved_totals")
        return M

    def stats_line(self, case_id, wall_s=0.0):
        buf = C.create_string_buffer(512)
        self._check(self.L.spb6_stats_line(self.h, int(case_id), float(wall_s), buf, 512), "stats_line")
        return buf.value.decode()

    @property
    def t(self):
        return self.L.spb6_time(self.h)

    # --- per-kernel access
    def residual(self, P):
        P = np.ascontiguousarray(P, dtype=np.float64).ravel()
        G = np.empty_like(P)
        bad = self.L.spb6_residual(self.h, _dp(P), _dp(G))
        if bad < 0:
            raise SPB6Error("residual failed")
        return G, bad

    def guess_residual(self, P):
        """learned guess at plain guess P after stage_prep: (the guess as a stage computes it, the same from the features
        of every cell, its residual as a stage computes it, the same residual evaluated on every cell, their two
        max-norms, the number of cells the first evaluated)"""
        P = np.ascontiguousarray(P, dtype=np.float64).ravel()
        Pg, Pgf, G, Gf, nrm = np.empty_like(P), np.empty_like(P), np.empty_like(P), np.empty_like(P), np.zeros(2)
        cnt = self.L.spb6_guess_residual(self.h, _dp(P), _dp(Pg), _dp(G), _dp(Gf), _dp(nrm), _dp(Pgf))
        if cnt < 0:
            raise SPB6Error("guess_residual failed (no network attached?)")
        return Pg, Pgf, G, Gf, nrm, cnt

    def jacobian(self, P, G):
        P = np.ascontiguousarray(P, dtype=np.float64).ravel()
        G = np.ascontiguousarray(G, dtype=np.float64).ravel()
        val = np.empty(5 * P.size)
        sc = np.empty(P.size)
        rc = self.L.spb6_jacobian(self.h, _dp(P), _dp(G), _dp(val), _dp(sc))
        if rc < 0:
            raise SPB6Error("jacobian failed")
        return val.reshape(-1, 5), sc, rc

    def linsolve(self, val, b, thomas=False):
        val = np.ascontiguousarray(val, dtype=np.float64).ravel()
        b = np.ascontiguousarray(b, dtype=np.float64).ravel()
        x = np.empty_like(b)
        it = C.c_int(0)
        rc = self.L.spb6_linsolve(self.h, _dp(val), _dp(b), _dp(x), int(thomas), C.byref(it))
        return x, it.value, rc

    def field(self, name):
        cap = self.NXT * self.NYT * 8
        out = np.empty(cap)
        n = self.L.spb6_get_field(self.h, name.encode(), _dp(out), cap)
        if n < 0:
            raise SPB6Error(f"unknown field {name}")
        return out[:n].copy()

    # timer categories, in the library's bit order (spb6_set_timing_mask)
    TIMING_CATEGORIES = ("explicit", "prep", "shock", "residual", "jacobian", "krylov", "final", "other", "krylov_aux", "guess",
                         "krylov_idle")

    def enable_timing(self, on=True, categories=None):
        """time the kernel categories (all, or the named ones: the others record no events) from now on, from zero"""
        mask = 0xFFFFFFFF if categories is None else sum(1 << self.TIMING_CATEGORIES.index(k) for k in categories)
        self._check(self.L.spb6_set_timing_mask(self.h, mask), "set_timing_mask")
        self._check(self.L.spb6_enable_timing(self.h, int(on)), "enable_timing")

    def kernel_times(self):
        t = KernelTimes()
        self._check(self.L.spb6_get_kernel_times(self.h, C.byref(t)), "kernel_times")
        return {k: getattr(t, k) for k, _ in KernelTimes._fields_}


