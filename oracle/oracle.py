"""ctypes wrapper of the CPU oracle (liboracle.so). TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg; never by the product package.
The oracle restates diffusionIMEX's SPB6_PC_IMEX / SPB6_PC_IMEX_omp time step on the CPU (see spb6_oracle.cpp).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

BC_TRANS, BC_WALL, BC_PERIODIC = 0, 1, 2


class Cfg(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int),
        ("x0", C.c_double), ("x1", C.c_double), ("y0", C.c_double), ("y1", C.c_double),
        ("bc", C.c_int * 4), ("bcu", C.c_double * 4),
        ("ph", (C.c_double * 6) * 2),
        ("rk", C.c_int), ("cfl", C.c_double), ("cfl_ac", C.c_double), ("ramp", C.c_int),
        ("dtfix", C.c_double), ("acdt", C.c_int),
        ("newton_tol", C.c_double), ("newton_iter", C.c_int),
        ("krylov_tol", C.c_double), ("krylov_iter", C.c_int),
        ("limiter", C.c_int), ("order", C.c_int), ("rus_local", C.c_int),
        ("closure_path", C.c_int), ("share_start", C.c_int), ("dot_order", C.c_int),
        ("portable_math", C.c_int), ("nthreads", C.c_int), ("jreuse", C.c_int), ("variant", C.c_int),
        ("relax", C.c_int), ("eps_i", C.c_double), ("shock_fused", C.c_int),
        ("theta_rate", C.c_double), ("nu_rate", C.c_double),
        ("eos_kind", C.c_int * 2), ("mg", (C.c_double * 7) * 2), ("mu_rate", C.c_double),
        ("nu_law", C.c_int), ("nu_cr", C.c_double), ("nu_pc", C.c_double),
        ("relax_form", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("steps", C.c_long), ("rk1", C.c_long), ("rejects", C.c_long), ("floor_acc", C.c_long),
        ("stages", C.c_long), ("newton_it", C.c_long), ("krylov_it", C.c_long), ("linsolves", C.c_long),
        ("maxres", C.c_double), ("t", C.c_double), ("dt_last", C.c_double), ("status", C.c_int),
        ("jac_builds", C.c_long),
    ]


class OraModel(C.Structure):
    _fields_ = [("ph", (C.c_double * 6) * 2), ("closure_path", C.c_int), ("share_start", C.c_int),
                ("dot_order", C.c_int), ("portable_math", C.c_int), ("shock_fused", C.c_int),
                ("eos_kind", C.c_int * 2), ("mg", (C.c_double * 7) * 2)]


def build(force=False):
    """Compile liboracle.so with the Makefile next to this file."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", HERE, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.ora_create.restype = vp
        L.ora_pexpm1.restype = C.c_double
        L.ora_pexpm1.argtypes = [C.c_double]
        L.ora_psat.restype = C.c_double
        L.ora_psat.argtypes = [vp, C.c_double]
        L.ora_nu_law.restype = C.c_double
        L.ora_nu_law.argtypes = [vp, C.c_double, C.c_double, C.c_double]
        L.ora_cell_nu.argtypes = [vp, dp]
        L.ora_closure_typed.argtypes = [C.POINTER(OraModel), dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, dp]
        L.ora_eos.argtypes = [C.POINTER(OraModel), C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, dp]
        L.ora_create.argtypes = [C.POINTER(Cfg)]
        L.ora_destroy.argtypes = [vp]
        L.ora_cfg_size.restype = C.c_size_t
        L.ora_stats_size.restype = C.c_size_t
        L.ora_set_state.argtypes = [vp, dp]
        L.ora_get_state.argtypes = [vp, dp]
        L.ora_get_pressure.argtypes = [vp, dp, dp]
        L.ora_run.argtypes = [vp, C.c_double, C.c_long]
        L.ora_get_stats.argtypes = [vp, C.POINTER(Stats)]
        L.ora_set_pressure.argtypes = [vp, dp, C.c_double]
        L.ora_set_clock.argtypes = [vp, C.c_double, C.c_double, C.c_int]
        L.ora_time.restype = C.c_double
        L.ora_time.argtypes = [vp]
        L.ora_stage_explicit.argtypes = [vp, C.c_double]
        L.ora_stage_prep.argtypes = [vp, C.c_double]
        L.ora_stage_newton.argtypes = [vp, ip, dp]
        L.ora_hllc_run.argtypes = [vp, C.c_double, C.c_long, C.c_double]
        L.ora_hllc_closure_fails.restype = C.c_long
        L.ora_hllc_step.argtypes = [vp, C.c_double]
        L.ora_relax_activate.argtypes = [vp]
        L.ora_tsat.restype = C.c_double
        L.ora_tsat.argtypes = [vp, C.c_double]
        L.ora_sat_info.argtypes = [vp, dp]
        L.ora_residual_all.argtypes = [vp, dp, dp]
        L.ora_jacobian.argtypes = [vp, dp, dp, dp, dp]
        L.ora_linsolve.argtypes = [vp, dp, dp, dp, C.c_int, ip]
        L.ora_field.restype = dp
        L.ora_field.argtypes = [vp, C.c_char_p]
        L.ora_pref.restype = C.c_double
        L.ora_pref.argtypes = [vp]
        L.ora_case_init.restype = C.c_double
        L.ora_case_init.argtypes = [C.c_int, C.POINTER(Cfg), dp]
        L.ora_cell_diag.argtypes = [C.POINTER(OraModel), dp, C.c_int, dp]
        L.ora_p_mix.restype = C.c_double
        L.ora_p_mix.argtypes = [C.POINTER(OraModel), C.c_double, C.c_double, C.c_double, C.c_double]
        L.ora_closure.argtypes = [C.POINTER(OraModel), dp, dp]
        L.ora_closure_phi.restype = C.c_double
        L.ora_closure_phi.argtypes = [C.POINTER(OraModel), dp, C.c_double]
        L.ora_shock_share.argtypes = [C.POINTER(OraModel), dp, dp, C.c_double, C.c_double, dp, dp]
        L.ora_hug_v.restype = C.c_double
        L.ora_hug_v.argtypes = [C.POINTER(OraModel), C.c_int, C.c_double, C.c_double, C.c_double]
        L.ora_prof_integrate.argtypes = [C.POINTER(OraModel), dp, dp, C.c_double, C.c_double, C.c_double, dp, dp]
        L.ora_prof_rhs.argtypes = [C.POINTER(OraModel), dp, dp, C.c_double, C.c_double, C.c_double, C.c_double, dp, dp]
        L.ora_shock_budget.argtypes = [C.POINTER(OraModel), dp, C.c_int, dp]
        for f in ("ora_pexp", "ora_plog"):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [C.c_double]
        L.ora_ppow.restype = C.c_double
        L.ora_ppow.argtypes = [C.c_double, C.c_double]
        L.ora_sum.restype = C.c_double
        L.ora_sum.argtypes = [C.c_int, dp, C.c_int]
        L.ora_set_threads.argtypes = [C.c_int]
        assert L.ora_cfg_size() == C.sizeof(Cfg), "Cfg layout mismatch"
        assert L.ora_stats_size() == C.sizeof(Stats), "Stats layout mismatch"
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def default_cfg(nx, ny=1, **kw):
    """Reference defaults (PC:57-72): LSDIRK2, material CFL 0.2, acoustic start-up CFL 0.4 with ramp, Newton tol 1e-13,
    60 Newton iterations, BiCGStab tol 1e-13 / 4000 iterations, minmod, local Rusanov speed."""
    c = Cfg()
    c.nx, c.ny = nx, ny
    c.x0, c.x1, c.y0, c.y1 = 0.0, 1.0, 0.0, 1.0
    c.rk, c.cfl, c.cfl_ac, c.ramp, c.dtfix, c.acdt = 2, 0.2, 0.4, 1, 0.0, 0
    c.newton_tol, c.newton_iter, c.krylov_tol, c.krylov_iter = 1e-13, 60, 1e-13, 4000
    c.limiter, c.order, c.rus_local = 1, 2, 1
    c.closure_path, c.share_start, c.dot_order, c.portable_math = 1, 1, 2, 1
    c.nthreads = 0
    c.jreuse = 0
    c.variant = 0
    c.relax, c.eps_i = 0, 1e-4
    c.shock_fused = 0
    c.theta_rate, c.nu_rate = -1.0, -1.0
    for k, v in kw.items():
        if k == "bc":
            for q in range(4):
                c.bc[q] = v[q]
        elif k == "bcu":
            for q in range(4):
                c.bcu[q] = v[q]
        elif k == "ph":
            for p in range(2):
                for q in range(6):
                    c.ph[p][q] = v[p][q]
        else:
            setattr(c, k, v)
    return c


def model_of(cfg):
    m = OraModel()
    for p in range(2):
        for q in range(6):
            m.ph[p][q] = cfg.ph[p][q]
    m.closure_path, m.share_start, m.dot_order, m.portable_math = cfg.closure_path, cfg.share_start, cfg.dot_order, cfg.portable_math
    m.shock_fused = cfg.shock_fused
    for p in range(2):
        m.eos_kind[p] = cfg.eos_kind[p]
        for q in range(7):
            m.mg[p][q] = cfg.mg[p][q]
    return m


def reference_mode(cfg, variant="PC"):
    """Switch a config to the reference's own iteration paths: "PC" (spb6_physics.h, B=4096 sums) or "PC_omp"
    (spb6_physics_omp.h, B=1024 sums), libm transcendental functions."""
    if variant == "PC":
        cfg.closure_path, cfg.share_start, cfg.dot_order = 0, 0, 0
    else:
        cfg.closure_path, cfg.share_start, cfg.dot_order = 1, 1, 1
    cfg.portable_math = 0
    return cfg


class Oracle:
    """One solver context (the reference's global grids and stage data)."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.L = lib()
        self.h = self.L.ora_create(C.byref(cfg))
        if not self.h:
            raise ValueError("a Mie-Grueneisen phase runs at relaxation level 0 with the reference scheme and the exact profile form only")
        self.nx, self.ny = cfg.nx, cfg.ny
        self.NXT, self.NYT = cfg.nx + 4, cfg.ny + 4

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.L.ora_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.ny, self.nx, 8)
        self.L.ora_set_state(self.h, _dp(U))

    def state(self):
        U = np.empty((self.ny, self.nx, 8))
        self.L.ora_get_state(self.h, _dp(U))
        return U

    def pressure(self):
        P = np.empty(self.nx * self.ny)
        pref = np.zeros(1)
        self.L.ora_get_pressure(self.h, _dp(P), _dp(pref))
        return P.reshape(self.ny, self.nx), float(pref[0])

    def set_pressure(self, P, pref):
        P = np.ascontiguousarray(P, dtype=np.float64).ravel()
        self.L.ora_set_pressure(self.h, _dp(P), float(pref))

    def set_clock(self, t, dtprev, startup):
        self.L.ora_set_clock(self.h, float(t), float(dtprev), int(bool(startup)))

    def run(self, t_end=1e300, max_steps=-1):
        return self.L.ora_run(self.h, float(t_end), int(max_steps))

    def hllc_run(self, t_end=1e300, max_steps=-1, cfl=0.4):
        """The reference's explicit HLLC scheme (src/SPB6_explicit.cpp) on this context's state, acoustic CFL cfl."""
        return self.L.ora_hllc_run(self.h, float(t_end), int(max_steps), float(cfl))

    def hllc_step(self, dt):
        """One step of the explicit HLLC scheme of size dt (0: accepted)."""
        return self.L.ora_hllc_step(self.h, float(dt))

    def stats(self):
        s = Stats()
        self.L.ora_get_stats(self.h, C.byref(s))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def field(self, name, grid=False):
        p = self.L.ora_field(self.h, name.encode())
        n = self.NXT * self.NYT * (8 if grid else 1)
        if name in ("P", "Pst"):
            n = self.nx * self.ny
        if name == "Qph":
            n = self.NXT * self.NYT * 2
        return np.ctypeslib.as_array(p, shape=(n,)).copy()


def case_init(test, nx, ny=1, **kw):
    """Reference test case `test` (stiffened-gas cases 0, 4, 5, 80-88 in 1D, 70 in 2D) on nx x ny cells:
    (cfg, U[ny * nx, 8], t_end)."""
    cfg = default_cfg(nx, ny, **kw)
    U = np.zeros((nx * ny, 8))
    tend = lib().ora_case_init(int(test), C.byref(cfg), _dp(U))
    if tend < 0:
        raise ValueError(f"unknown case {test}")
    for k, v in kw.items():   # explicit overrides win over the case defaults
        if k not in ("bc", "bcu", "ph", "eos_kind", "mg"):
            setattr(cfg, k, v)
    return cfg, U, tend


def cell_diag(cfg, U):
    """Profile columns p u v rho alpha2 Y2 rho1 rho2 c_frozen c_wood T1 T2 p1 p2 (PHY:1105-1116)."""
    U = np.ascontiguousarray(U, dtype=np.float64).reshape(-1, 8)
    out = np.empty((U.shape[0], 14))
    m = model_of(cfg)
    lib().ora_cell_diag(C.byref(m), _dp(U), U.shape[0], _dp(out))
    return out
