"""Modified Newton (jreuse = 1, src/SPB6_PC_IMEX_omp.cpp:796-888) on the GPU against the oracle.

Plain form: whole runs bit for bit, with the same Jacobian builds, Newton iterations and Krylov iterations. Mixed form
(FP32 BiCGStab inside the FP64 Newton; a reused Jacobian keeps its FP32 copy): within the Newton tolerance's reach of
the FP64 oracle run with jreuse, relative L1 1e-8 per field on the cavitation flow.
"""
import pytest

from tests.helpers import bits_diff, oracle_cfg_from, field_l1

pytestmark = pytest.mark.gpu


def _cases():
    from diffusionimex_amd import cases
    from diffusionimex_amd.spb6 import BC_PERIODIC
    return [
        ("tube200", cases.water_air_tube(200, cfl=0.5, newton_tol=1e-10, jreuse=1), 40),
        ("bubble48x24", cases.shock_bubble(48, 24, seed=3, cfl=0.5, jreuse=1), 20),
        ("cavitation33x21_periodic", cases.cavitation_2d(33, 21, seed=4, cfl=0.5, bc=(BC_PERIODIC,) * 4, jreuse=1), 15),
    ]


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_jreuse_run_bitwise(idx):
    from diffusionimex_amd import spb6
    from oracle import oracle as O
    name, (cfg, U), steps = _cases()[idx]
    assert cfg.jreuse == 1
    o = O.Oracle(oracle_cfg_from(cfg, O))
    o.set_state(U)
    assert o.run(1.0, steps) == 0
    st = spb6.Stepper(cfg)
    st.set_state(U)
    assert st.run(1.0, steps) == 0
    so, ss = o.stats(), st.stats()
    for k in ("steps", "rk1", "rejects", "stages", "newton_it", "krylov_it", "floor_acc", "linsolves", "jac_builds"):
        assert so[k] == ss[k], f"{name}: stat {k} {so[k]} vs {ss[k]}"
    assert ss["jac_builds"] < ss["linsolves"], f"{name}: no Jacobian was reused"
    Uo, Us = o.state(), st.state()
    assert bits_diff(Us, Uo) == 0, f"{name}: state differs in {bits_diff(Us, Uo)} values; per-field L1 {field_l1(Us, Uo)}"
    assert bits_diff(st.pressure()[0], o.pressure()[0]) == 0


def test_jreuse_mixed_within_1e8():
    from diffusionimex_amd import cases, spb6
    from oracle import oracle as O
    cfg, U = cases.cavitation_2d(64, 64, seed=0, jreuse=1)
    o = O.Oracle(oracle_cfg_from(cfg, O))
    o.set_state(U)
    assert o.run(1.0, 30) == 0
    st = spb6.Stepper(cases.cavitation_2d(64, 64, seed=0, jreuse=1, krylov_fp32=1)[0])
    st.set_state(U)
    assert st.run(1.0, 30) == 0
    ss = st.stats()
    assert ss["steps"] == o.stats()["steps"] and ss["jac_builds"] < ss["linsolves"]
    l1 = field_l1(st.state(), o.state())
    assert max(l1) <= 1e-8, f"per-field L1 {l1}"
