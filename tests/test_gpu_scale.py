"""GPU parity at the sizes where the launch structure changes, at the small workload's own configuration, and on the
stepper's failure paths -- each against the CPU oracle in its device-order mode, bit for bit.

- Multi-block Krylov reductions: a 700 x 300 grid gives 206 FP64 BiCGStab blocks (206 mod 8 = 6, so the XCD bands are
  uneven) whose partials every pass re-reduces; 2050 x 1026 rows (above 256 * 4 * 2048) switch the passes to 8 rows per
  thread, which changes the blocks of the summation tree (the oracle's dev_items follows the same rule).
- The shock budget's strong-shock pass on a cavitating flow, where most compressive cells need the RK45 shooting.
- The 1000-cell water-air tube at CFL 0.5 and Newton tolerance 1e-10 for its full 200 steps.
- Failure paths of the time loop (PC:1240, PC:1287-1290): Newton that cannot reach its tolerance (rejects, then the
  step fails after 30 halvings or is taken at the round-off floor), a state with a negative density (the explicit
  stage's check, PC:257), and dt below 1e-12 t_end. Each ends with a host-visible status code.
"""
import numpy as np
import pytest

from tests.helpers import bits_diff, oracle_cfg_from, field_l1

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def O():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="module")
def S():
    from diffusionimex_amd import spb6
    spb6.load()
    return spb6


STAT_KEYS = ("steps", "rk1", "rejects", "stages", "newton_it", "krylov_it", "floor_acc")


def _run_both(S, O, cfg, U, steps, t_end=1.0):
    o = O.Oracle(oracle_cfg_from(cfg, O))
    o.set_state(U)
    st = S.Stepper(cfg)
    st.set_state(U)
    rco = o.run(t_end, steps)
    rcs = st.run(t_end, steps)
    return o, st, rco, rcs


def _assert_same(name, o, st):
    so, ss = o.stats(), st.stats()
    for k in STAT_KEYS:
        assert so[k] == ss[k], f"{name}: stat {k} {so[k]} vs {ss[k]}"
    Uo, Us = o.state(), st.state()
    assert bits_diff(Us, Uo) == 0, f"{name}: state differs in {bits_diff(Us, Uo)} values; per-field L1 {field_l1(Us, Uo)}"
    Po, _ = o.pressure()
    Ps, _ = st.pressure()
    assert bits_diff(Ps, Po) == 0, f"{name}: pressure differs in {bits_diff(Ps, Po)} cells"


def test_plain_run_bitwise_multiblock(S, O):
    """206 Krylov blocks with uneven XCD bands, 22 x 38 explicit-stage tiles: 20 steps (the 17 first-order start-up
    steps and 3 LSDIRK2 steps) bit for bit"""
    from diffusionimex_amd import cases
    cfg, U = cases.shock_bubble(700, 300, seed=1, cfl=0.5)
    o, st, rco, rcs = _run_both(S, O, cfg, U, 20)
    assert rco == 0 and rcs == 0
    assert st.stats()["krylov_it"] > 0
    _assert_same("bubble700x300", o, st)


def test_lsdirk2_step_bitwise_eight_rows_per_thread(S, O):
    """2050 x 1026 = 2,103,300 rows: one full LSDIRK2 step (no start-up ramp) with the 8-rows-per-thread Krylov tree"""
    from diffusionimex_amd import cases
    cfg, U = cases.shock_bubble(2050, 1026, seed=2, cfl=0.5, ramp=0)
    assert cfg.nx * cfg.ny > 256 * 4 * 2048
    o, st, rco, rcs = _run_both(S, O, cfg, U, 1)
    assert rco == 0 and rcs == 0
    assert st.stats()["rk1"] == 0 and st.stats()["stages"] == 2
    _assert_same("bubble2050x1026", o, st)


def test_shock_budget_cavitation_bitwise(S, O):
    """the strong-shock pass where most compressive cells need it: a 160 x 128 cavitating flow after 20 steps"""
    from diffusionimex_amd import cases
    cfg, U = cases.cavitation_2d(160, 128, seed=5, cfl=0.5)
    o = O.Oracle(oracle_cfg_from(cfg, O))
    o.set_state(U)
    assert o.run(1.0, 20) == 0
    st = S.Stepper(cfg)
    st.set_state(o.state())
    P, pref = o.pressure()
    st.set_pressure(P, pref)
    adt = 0.29 * o.stats()["dt_last"]
    assert o.L.ora_stage_explicit(o.h, adt) == 0 and st.explicit_stage(adt) == 0
    assert o.L.ora_stage_prep(o.h, adt) == 0 and st.stage_prep(adt) == 0
    qa = st.field("Qph").reshape(2, st.NYT, st.NXT)[:, 2:-2, 2:-2]
    qb = o.field("Qph").reshape(st.NYT, st.NXT, 2)[2:-2, 2:-2, :].transpose(2, 0, 1)
    assert np.count_nonzero(qb[0]) > 0.2 * cfg.nx * cfg.ny
    assert bits_diff(qa, qb) == 0, f"shock budget differs in {bits_diff(qa, qb)} values"


def test_water_air_tube_c0_full_run(S, O):
    """the 1000-cell water-air tube at its own settings: CFL 0.5, Newton tolerance 1e-10, 200 steps"""
    from diffusionimex_amd import cases
    cfg, U = cases.water_air_tube(1000, cfl=0.5, newton_tol=1e-10)
    o, st, rco, rcs = _run_both(S, O, cfg, U, 200)
    assert rco == 0 and rcs == 0
    assert st.stats()["steps"] == 200
    _assert_same("tube1000", o, st)


def test_unreachable_newton_tolerance(S, O):
    """Newton tolerance 1e-20: every stage's Newton stops at the round-off floor or fails; the floor acceptances,
    rejects, halvings and the final status are the oracle's"""
    from diffusionimex_amd import cases
    cfg, U = cases.water_air_tube(100, cfl=0.5, newton_tol=1e-20, newton_iter=8)
    o, st, rco, rcs = _run_both(S, O, cfg, U, 3)
    assert (rco == 0) == (rcs == 0)
    so, ss = o.stats(), st.stats()
    assert so["floor_acc"] + so["rejects"] > 0
    if rco != 0:
        assert rcs == S.STATUS_CODE["STEP_FAILED"] and so["status"] == 2
    _assert_same("tube100_tol1e-20", o, st)


def test_negative_density_rejects_then_fails(S, O):
    """a cell with m1 < 0 in an otherwise uniform gas: the explicit stage's density check (PC:257) rejects every
    attempt, dt halves 30 times and the step fails (PC:1290), as in the oracle"""
    from diffusionimex_amd import cases
    cfg, U = cases.water_air_tube(64, cfl=0.5)
    U = U.copy()
    U[0, 40, 1] = -abs(U[0, 40, 1])   # m1 < 0: rho = m1 + m2 < 0 in this water cell
    U[0, 40, 2] = 0.0
    o, st, rco, rcs = _run_both(S, O, cfg, U, 1)
    so, ss = o.stats(), st.stats()
    assert rco != 0 and rcs != 0
    expect = {2: "STEP_FAILED", 3: "DT_UNDERFLOW"}[so["status"]]
    assert rcs == S.STATUS_CODE[expect], f"status {S.STATUS.get(rcs)} vs oracle {expect}"
    for k in ("steps", "rejects", "rk1"):
        assert so[k] == ss[k], f"stat {k} {so[k]} vs {ss[k]}"


def test_dt_underflow(S, O):
    """dt below 1e-12 t_end (PC:1240): with t_end = 1e8 s the first CFL step of the tube (~2e-6 s) is far below 1e-4 s"""
    from diffusionimex_amd import cases
    cfg, U = cases.water_air_tube(100, cfl=0.5)
    o, st, rco, rcs = _run_both(S, O, cfg, U, 5, t_end=1e8)
    assert rco == 2 and o.stats()["status"] == 3
    assert rcs == S.STATUS_CODE["DT_UNDERFLOW"]
    assert st.stats()["steps"] == 0 and st.stats()["status"] == S.STATUS_CODE["DT_UNDERFLOW"]


@pytest.mark.parametrize("which", ["tube", "bubble"])
def test_ssp2_332_table_bitwise(S, O, which):
    """rk = 3, the reference's SSP2(3,3,2) table (PC:1206-1216): three stages per step, the third combining both earlier
    stage derivatives, bit for bit against the oracle"""
    from diffusionimex_amd import cases
    if which == "tube":
        cfg, U = cases.water_air_tube(200, cfl=0.2, newton_tol=1e-13, rk=3)
        steps = 40
    else:
        cfg, U = cases.shock_bubble(48, 24, seed=3, cfl=0.5, rk=3)
        steps = 25
    o, st, rco, rcs = _run_both(S, O, cfg, U, steps)
    assert rco == 0 and rcs == 0
    s = st.stats()
    assert s["rejects"] == 0 and s["steps"] > s["rk1"] and s["stages"] == s["rk1"] + 3 * (s["steps"] - s["rk1"])
    _assert_same(f"{which}_rk3", o, st)
