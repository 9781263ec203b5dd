"""The reference's IMEX variants (SPB6_TV / TVpure / TVinc / TVsplit / TVbb / BB_IMEX.cpp) on the GPU against the oracle,
bit for bit: whole runs (state, pressure, step, Newton, Krylov and floor counters) in the oracle's device mode, which
in the reference's own mode reproduces the reference's results for these variants (test_oracle_variants)."""
import pytest

from tests.helpers import bits_diff, oracle_cfg_from, field_l1

pytestmark = pytest.mark.gpu

VARIANTS = ["TV", "TVpure", "TVinc", "TVsplit", "TVbb", "BB"]


def _case(name, v):
    from diffusionimex_amd import cases
    from diffusionimex_amd.spb6 import BC_PERIODIC
    if name == "tube":
        return cases.water_air_tube(200, cfl=0.5, newton_tol=1e-10, variant=v), 30
    if name == "bubble":
        return cases.shock_bubble(48, 24, seed=3, cfl=0.5, variant=v), 12
    if name == "interface":   # odd sizes, walls in y
        return cases.perturbed_interface(37, 29, seed=2, cfl=0.5, variant=v), 10
    return cases.cavitation_2d(33, 21, seed=4, cfl=0.5, bc=(BC_PERIODIC,) * 4, variant=v), 10


@pytest.mark.parametrize("case", ["tube", "bubble", "interface", "cavitation"])
@pytest.mark.parametrize("var", VARIANTS)
def test_variant_run_bitwise(var, case):
    from diffusionimex_amd import spb6
    from oracle import oracle as O
    v = spb6.VARIANTS[var]
    (cfg, U), steps = _case(case, v)
    assert cfg.variant == v
    o = O.Oracle(oracle_cfg_from(cfg, O))
    o.set_state(U)
    rco = o.run(1.0, steps)
    st = spb6.Stepper(cfg)
    st.set_state(U)
    rcs = st.run(1.0, steps)
    assert rco == 0 and rcs == 0, (var, case, rco, rcs)
    so, ss = o.stats(), st.stats()
    for k in ("steps", "rk1", "rejects", "stages", "newton_it", "krylov_it", "floor_acc", "linsolves"):
        assert so[k] == ss[k], f"{var}/{case}: {k} {so[k]} vs {ss[k]}"
    Uo, Us = o.state(), st.state()
    assert bits_diff(Us, Uo) == 0, f"{var}/{case}: state differs in {bits_diff(Us, Uo)} values; per-field L1 {field_l1(Us, Uo)}"
    assert bits_diff(st.pressure()[0], o.pressure()[0]) == 0


def test_variant_rejects_fast_forms():
    from diffusionimex_amd import cases, spb6
    cfg, _ = cases.water_air_tube(50, variant=1, analytic_jacobian=1)
    with pytest.raises(spb6.SPB6Error):
        spb6.Stepper(cfg)
