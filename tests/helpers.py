"""Shared test helpers: oracle <-> stepper configuration and comparison utilities."""
import numpy as np

from diffusionimex_amd.spb6 import make_config


def config_from_oracle(ocfg, **over):
    """Stepper Config equivalent to an oracle Cfg."""
    c = make_config(ocfg.nx, ocfg.ny,
                    phases=[tuple(ocfg.ph[k][q] for q in range(6)) for k in range(2)],
                    domain=(ocfg.x0, ocfg.x1, ocfg.y0, ocfg.y1), bc=tuple(ocfg.bc), bcu=tuple(ocfg.bcu),
                    rk=ocfg.rk, cfl=ocfg.cfl, cfl_ac=ocfg.cfl_ac, ramp=ocfg.ramp, dtfix=ocfg.dtfix, acdt=ocfg.acdt,
                    newton_tol=ocfg.newton_tol, newton_iter=ocfg.newton_iter, krylov_tol=ocfg.krylov_tol,
                    krylov_iter=ocfg.krylov_iter, limiter=ocfg.limiter, order=ocfg.order, rus_local=ocfg.rus_local,
                    jreuse=ocfg.jreuse, variant=ocfg.variant, relax=ocfg.relax, eps_i=ocfg.eps_i,
                    shock_fused=ocfg.shock_fused, theta_rate=max(ocfg.theta_rate, 0.0), nu_rate=max(ocfg.nu_rate, 0.0),
                    eos_kind=tuple(ocfg.eos_kind), mg=[tuple(ocfg.mg[k][q] for q in range(7)) for k in range(2)],
                    mu_rate=max(ocfg.mu_rate, 0.0), nu_law=ocfg.nu_law, nu_cr=ocfg.nu_cr, nu_pc=ocfg.nu_pc,
                    relax_form=ocfg.relax_form)
    for k, v in over.items():
        setattr(c, k, v)
    return c


def oracle_cfg_from(scfg, O, **over):
    """Oracle Cfg equivalent to a stepper Config (device-order sums, PC_omp closure path, portable math)."""
    ph = [[scfg.ph[k].gamma, scfg.ph[k].pi_inf, scfg.ph[k].q, scfg.ph[k].qp, scfg.ph[k].cv, scfg.ph[k].b] for k in range(2)]
    c = O.default_cfg(scfg.nx, scfg.ny, bc=list(scfg.bc), bcu=list(scfg.bcu), ph=ph)
    c.x0, c.x1, c.y0, c.y1 = scfg.x0, scfg.x1, scfg.y0, scfg.y1
    for k in ("rk", "cfl", "cfl_ac", "ramp", "dtfix", "acdt", "newton_tol", "newton_iter", "krylov_tol", "krylov_iter",
              "limiter", "order", "rus_local", "jreuse", "variant", "relax", "shock_fused"):
        setattr(c, k, getattr(scfg, k))
    c.eps_i = scfg.eps_i if scfg.eps_i > 0 else 1e-4
    c.theta_rate = scfg.theta_rate if scfg.theta_rate > 0 else -1.0
    c.nu_rate = scfg.nu_rate if scfg.nu_rate > 0 else -1.0
    c.mu_rate = scfg.mu_rate if scfg.mu_rate > 0 else -1.0
    c.nu_law, c.nu_cr, c.nu_pc = scfg.nu_law, scfg.nu_cr, scfg.nu_pc
    c.relax_form = scfg.relax_form
    for k in range(2):
        c.eos_kind[k] = scfg.eos_kind[k]
        for q in range(7):
            c.mg[k][q] = scfg.mg[k][q]
    for k, v in over.items():
        setattr(c, k, v)
    return c


def bits_diff(a, b):
    """Number of elements whose IEEE bits differ (NaN == NaN)."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    both_nan = np.isnan(a) & np.isnan(b)
    return int(np.count_nonzero((a.view(np.uint64) != b.view(np.uint64)) & ~both_nan))


def rel_l1(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


def field_l1(Ua, Ub):
    """per-field relative L1 of two [.., 8] state arrays"""
    Ua = Ua.reshape(-1, 8)
    Ub = Ub.reshape(-1, 8)
    return [rel_l1(Ua[:, k], Ub[:, k]) for k in range(8)]
