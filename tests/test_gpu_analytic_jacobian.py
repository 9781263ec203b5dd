"""Analytic Newton Jacobian (config analytic_jacobian = 1): each row from the implicit derivative of the closure at the
residual's root, instead of the reference's five finite-difference closure solves per row (PC:790-804).

The entries are checked against the oracle's finite-difference Jacobian of the same state (the reference's own
algorithm), row-scaled: a forward difference with step 1e-7 s carries relative errors of about 1e-7 (truncation) plus
the closure's stopping tolerance divided by the step, so the bar is 1e-5 of the row's largest entry. Runs with the
option converge to the same Newton tolerance; on well-conditioned flows every field stays within relative L1 1e-8 of the
FP64 oracle, on the shock-bubble within the scheme's own spread (as for the FP32 Krylov form)."""
import numpy as np
import pytest

from tests.helpers import oracle_cfg_from, field_l1

pytestmark = pytest.mark.gpu

ROW_TOL = 1e-5   # |analytic - finite difference| / max |entry| of the row
L1_TOL = 1e-8    # per-field relative L1 against the FP64 oracle


def _cases():
    from diffusionimex_amd import cases
    out = []
    cfg, U = cases.water_air_tube(200, cfl=0.2, newton_tol=1e-13)
    out.append(("tube200", cfg, U, 30))
    cfg, U = cases.shock_bubble(48, 24, seed=3, cfl=0.5)
    out.append(("bubble48x24", cfg, U, 8))
    cfg, U = cases.cavitation_2d(64, 64, seed=0)
    out.append(("cavitation64", cfg, U, 10))
    return out


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_analytic_jacobian_matches_oracle_finite_differences(idx):
    from diffusionimex_amd import spb6
    from oracle import oracle as O
    name, cfg, U, warm = _cases()[idx]
    o = O.Oracle(oracle_cfg_from(cfg, O))
    o.set_state(U)
    assert o.run(1.0, warm) == 0
    cfg.analytic_jacobian = 1
    st = spb6.Stepper(cfg)
    st.set_state(o.state())
    P, pref = o.pressure()
    st.set_pressure(P, pref)
    adt = 0.29 * o.stats()["dt_last"]
    assert o.L.ora_stage_explicit(o.h, adt) == 0 and st.explicit_stage(adt) == 0
    assert o.L.ora_stage_prep(o.h, adt) == 0 and st.stage_prep(adt) == 0
    n = cfg.nx * cfg.ny
    P = P.ravel().copy()
    Go = np.empty(n)
    assert o.L.ora_residual_all(o.h, O._dp(P), O._dp(Go)) == 0
    Gs, bad = st.residual(P)
    assert bad == 0
    vo = np.empty(5 * n)
    so = np.empty(n)
    assert o.L.ora_jacobian(o.h, O._dp(P), O._dp(Go), O._dp(vo), O._dp(so)) == 0
    va, sa, rc = st.jacobian(P, Gs)
    assert rc == 0
    fd = vo.reshape(n, 5)
    ncol = 5 if cfg.ny > 1 else 3
    fd, an = fd[:, :ncol], va[:, :ncol]
    scale = np.abs(fd).max(axis=1, keepdims=True)
    err = np.abs(an - fd) / scale
    assert np.isfinite(an).all()
    assert err.max() <= ROW_TOL, f"{name}: worst row-scaled difference {err.max():.2e} at row {np.unravel_index(err.argmax(), err.shape)}"
    # the diagonal carries the -1 of G = pbar - P and dominates: same sign as the reference's
    assert (np.sign(an[:, 0]) == np.sign(fd[:, 0])).all()


def _run_pair(mk, U, steps, **kw):
    from diffusionimex_amd import spb6
    from oracle import oracle as O
    o = O.Oracle(oracle_cfg_from(mk(), O))
    o.set_state(U)
    assert o.run(1.0, steps) == 0
    st = spb6.Stepper(mk(analytic_jacobian=1, **kw))
    st.set_state(U)
    assert st.run(1.0, steps) == 0
    return o, st


@pytest.mark.parametrize("fp32", [0, 1])
def test_analytic_jacobian_cavitation_within_1e8_of_fp64_oracle(fp32):
    from diffusionimex_amd import cases
    _, U = cases.cavitation_2d(64, 64, seed=0)
    mk = lambda **kw: cases.cavitation_2d(64, 64, seed=0, **kw)[0]
    o, st = _run_pair(mk, U, 30, krylov_fp32=fp32)
    so, ss = o.stats(), st.stats()
    assert ss["steps"] == so["steps"] and ss["rejects"] == 0
    assert ss["maxres"] < 1e-13
    assert ss["newton_it"] <= so["newton_it"]
    l1 = field_l1(st.state(), o.state())
    assert max(l1) <= L1_TOL, f"per-field L1 {l1}"


def test_analytic_jacobian_shock_bubble_within_scheme_spread():
    from diffusionimex_amd import cases
    from oracle import oracle as O
    mk = lambda **kw: cases.shock_bubble(48, 24, seed=1, cfl=0.5, **kw)[0]
    _, U = cases.shock_bubble(48, 24, seed=1, cfl=0.5)
    o, st = _run_pair(mk, U, 25, krylov_fp32=1)
    ss = st.stats()
    assert ss["rejects"] == 0 and ss["maxres"] < 1e-13
    o2 = O.Oracle(oracle_cfg_from(mk(), O, krylov_tol=1e-12))
    o2.set_state(U)
    assert o2.run(1.0, 25) == 0
    spread = field_l1(o2.state(), o.state())
    l1 = field_l1(st.state(), o.state())
    for k in range(8):
        assert l1[k] <= 10 * spread[k] + 1e-12, f"field {k}: {l1[k]:.2e} vs spread {spread[k]:.2e}"
