"""GPU parity: every kernel of the HIP stepper against the CPU oracle on the same inputs, bit for bit.

The oracle runs in its device-order mode (PC_omp iteration paths, the HIP kernels' fixed reduction order, portable
exp/log/pow), which is itself pinned to the reference's archived outputs in test_oracle_fixtures.py.
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


def _pair(S, O, scfg, U, warm_steps=0):
    """oracle + stepper on the same state; the oracle first advances warm_steps steps and the stepper takes its state."""
    ocfg = oracle_cfg_from(scfg, O)
    o = O.Oracle(ocfg)
    o.set_state(U)
    if warm_steps:
        assert o.run(1.0, warm_steps) == 0
    st = S.Stepper(scfg)
    st.set_state(o.state())
    P, pref = o.pressure()
    st.set_pressure(P, pref)
    return o, st


def _cases():
    from diffusionimex_amd import cases
    out = []
    cfg, U = cases.water_air_tube(200, cfl=0.2, newton_tol=1e-13)
    out.append(("tube200", cfg, U, 30))
    cfg, U = cases.shock_bubble(48, 24, seed=3, cfl=0.5)
    out.append(("bubble48x24", cfg, U, 8))
    # odd sizes: the explicit stage's 32 x 8 tiles end part-way in both directions (ghosted 41 x 33, 37 x 25)
    cfg, U = cases.perturbed_interface(37, 29, seed=2, cfl=0.5)
    out.append(("interface37x29", cfg, U, 6))
    from diffusionimex_amd.spb6 import BC_PERIODIC
    cfg, U = cases.cavitation_2d(33, 21, seed=4, cfl=0.5, bc=(BC_PERIODIC,) * 4)
    out.append(("cavitation33x21_periodic", cfg, U, 6))
    return out




@pytest.mark.parametrize("idx", [0, 1, 2, 3])
def test_explicit_prep_shock_bitwise(S, O, idx):
    name, cfg, U, warm = _cases()[idx]
    o, st = _pair(S, O, cfg, U, warm)
    adt = 0.29 * o.stats()["dt_last"]
    assert o.L.ora_stage_explicit(o.h, adt) == 0
    assert st.explicit_stage(adt) == 0
    NC = st.NXT * st.NYT
    for f in ["Ust", "wxL", "wxR"] + (["wyL", "wyR"] if cfg.ny > 1 else []):
        a, b = st.field(f).reshape(NC, 8), o.field(f, grid=True).reshape(NC, 8)
        assert bits_diff(a, b) == 0, f"{name}: field {f} differs in {bits_diff(a, b)} values"
    for f in ["a2s", "phx", "ubx"] + (["phy", "uby"] if cfg.ny > 1 else []):
        a, b = st.field(f), o.field(f)
        if f in ("ubx", "uby"):   # faces are defined on interior rows only
            a = a.reshape(st.NYT, st.NXT)[2:-2, 1:-2] if f == "ubx" else a.reshape(st.NYT, st.NXT)[1:-2, 2:-2]
            b = b.reshape(st.NYT, st.NXT)[2:-2, 1:-2] if f == "ubx" else b.reshape(st.NYT, st.NXT)[1:-2, 2:-2]
        assert bits_diff(a, b) == 0, f"{name}: field {f} differs in {bits_diff(a, b)} values"
    assert o.L.ora_stage_prep(o.h, adt) == 0
    assert st.stage_prep(adt) == 0
    for f in ["ust", "vst", "cz", "Kw", "ubsx", "kpx", "wLx", "pvx"] + (["ubsy", "kpy", "wLy", "pvy"] if cfg.ny > 1 else []):
        a, b = st.field(f).reshape(st.NYT, st.NXT), o.field(f).reshape(st.NYT, st.NXT)
        if f in ("ubsx", "kpx", "wLx", "pvx"):
            a, b = a[:, :-1], b[:, :-1]
        if f in ("ubsy", "kpy", "wLy", "pvy"):
            a, b = a[:-1, :], b[:-1, :]
        if cfg.ny == 1:
            a, b = a[2:3], b[2:3]
        assert bits_diff(a, b) == 0, f"{name}: prep field {f} differs in {bits_diff(a, b)} values"
    qa = st.field("Qph").reshape(2, st.NYT, st.NXT)[:, 2:-2, 2:-2]
    qb = o.field("Qph").reshape(st.NYT, st.NXT, 2)[2:-2, 2:-2, :].transpose(2, 0, 1)
    assert np.count_nonzero(qb) > 0 or idx != 0
    assert bits_diff(qa, qb) == 0, f"{name}: shock budget differs in {bits_diff(qa, qb)} values"


@pytest.mark.parametrize("idx", [0, 1, 2, 3])
def test_residual_jacobian_linsolve_bitwise(S, O, idx):
    name, cfg, U, warm = _cases()[idx]
    o, st = _pair(S, O, cfg, U, warm)
    adt = 0.29 * o.stats()["dt_last"]
    assert o.L.ora_stage_explicit(o.h, adt) == 0 and st.explicit_stage(adt) == 0
    assert o.L.ora_stage_prep(o.h, adt) == 0 and st.stage_prep(adt) == 0
    n = cfg.nx * cfg.ny
    P, _ = o.pressure()
    P = P.ravel().copy()
    Go = np.empty(n)
    bad = o.L.ora_residual_all(o.h, O._dp(P), O._dp(Go))
    Gs, bad_s = st.residual(P)
    assert bad == bad_s == 0
    assert bits_diff(Gs, Go) == 0, f"{name}: residual differs in {bits_diff(Gs, Go)} cells"
    vo = np.empty(5 * n)
    so = np.empty(n)
    assert o.L.ora_jacobian(o.h, O._dp(P), O._dp(Go), O._dp(vo), O._dp(so)) == 0
    vs, ss, rc = st.jacobian(P, Gs)
    assert rc == 0
    assert bits_diff(ss, so) == 0
    assert bits_diff(vs.ravel(), vo) == 0, f"{name}: Jacobian differs in {bits_diff(vs.ravel(), vo)} entries"
    rhs = -Go
    thomas = cfg.ny == 1
    xo = np.empty(n)
    ito = np.zeros(1, dtype=np.int32)
    import ctypes
    rco = o.L.ora_linsolve(o.h, O._dp(vo), O._dp(rhs), O._dp(xo), int(thomas), ito.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    xs, its, rcs = st.linsolve(vo, rhs, thomas=thomas)
    assert rco == rcs == 0
    assert its == int(ito[0])
    assert bits_diff(xs, xo) == 0, f"{name}: linear solve differs in {bits_diff(xs, xo)} entries"


@pytest.mark.parametrize("idx,steps", [(0, 60), (1, 25), (2, 20), (3, 20)])
def test_run_bitwise(S, O, idx, steps):
    name, cfg, U, _ = _cases()[idx]
    o, st = _pair(S, O, cfg, U, 0)
    assert o.run(1.0, steps) == 0
    assert st.run(1.0, steps) == 0
    so, ss = o.stats(), st.stats()
    for k in ("steps", "rk1", "rejects", "stages", "newton_it", "krylov_it", "floor_acc"):
        assert so[k] == ss[k], f"{name}: stat {k} {so[k]} vs {ss[k]}"
    Uo, Us = o.state(), st.state()
    assert bits_diff(Us, Uo) == 0, f"{name}: state differs in {bits_diff(Us, Uo)} values; per-field L1 {field_l1(Us, Uo)}"
    Po, _ = o.pressure()
    Ps, _ = st.pressure()
    assert bits_diff(Ps, Po) == 0
