"""The strong-shock pass in fused form (shock_fused = 1: prof_step_fused in k_shock_strong) on the GPU against the
oracle's restatement of that form (prof_integrate_fused), bit for bit: whole runs of the water-air tube, a 2D
shock-bubble, a perturbed interface on odd sizes and a periodic cavitation box, with the plain and the mixed forms
(the fused shock pass is part of the bench's mixed and learned forms). The fused form itself is held to the reference's
arithmetic within round-off by tests/test_oracle_shock_fused.py."""
import pytest

from tests.helpers import bits_diff, field_l1, oracle_cfg_from

pytestmark = pytest.mark.gpu


def _case(name):
    from diffusionimex_amd import cases
    from diffusionimex_amd.spb6 import BC_PERIODIC
    if name == "tube":
        return cases.water_air_tube(200, cfl=0.5, newton_tol=1e-10, shock_fused=1), 40
    if name == "bubble":
        return cases.shock_bubble(48, 24, seed=3, cfl=0.5, shock_fused=1), 12
    if name == "interface":
        return cases.perturbed_interface(37, 29, seed=2, cfl=0.5, shock_fused=1), 10
    return cases.cavitation_2d(33, 21, seed=4, cfl=0.5, bc=(BC_PERIODIC,) * 4, shock_fused=1), 10


@pytest.mark.parametrize("case", ["tube", "bubble", "interface", "cavitation"])
def test_fused_shock_run_bitwise(case):
    from diffusionimex_amd import spb6
    from oracle import oracle as O
    (cfg, U), steps = _case(case)
    assert cfg.shock_fused == 1
    oc = oracle_cfg_from(cfg, O)
    assert oc.shock_fused == 1
    o = O.Oracle(oc)
    o.set_state(U)
    rco = o.run(1.0, steps)
    st = spb6.Stepper(cfg)
    st.set_state(U)
    rcs = st.run(1.0, steps)
    assert rco == 0 and rcs == 0, (case, rco, rcs)
    so, ss = o.stats(), st.stats()
    for k in ("steps", "rk1", "rejects", "stages", "newton_it", "krylov_it", "floor_acc"):
        assert so[k] == ss[k], f"{case}: {k} {so[k]} vs {ss[k]}"
    Uo, Us = o.state(), st.state()
    assert bits_diff(Us, Uo) == 0, f"{case}: state differs in {bits_diff(Us, Uo)} values; per-field L1 {field_l1(Us, Uo)}"


def test_fused_shock_pass_on_a_large_list():
    """a 512 x 256 cavitation box (tens of thousands of strong cells, every lane of many wavefronts busy) for two
    steps, fused, against the oracle bit for bit"""
    from diffusionimex_amd import cases, spb6
    from diffusionimex_amd.spb6 import BC_PERIODIC
    from oracle import oracle as O
    cfg, U = cases.cavitation_2d(512, 256, seed=1, cfl=0.5, bc=(BC_PERIODIC,) * 4, shock_fused=1)
    o = O.Oracle(oracle_cfg_from(cfg, O))
    o.set_state(U)
    assert o.run(1.0, 2) == 0
    st = spb6.Stepper(cfg)
    st.set_state(U)
    assert st.run(1.0, 2) == 0
    assert bits_diff(st.state(), o.state()) == 0


def test_batch_rejects_fused():
    """the batched 1D stepper runs the reference's profile arithmetic only"""
    from diffusionimex_amd import batch, spb6
    scheme = spb6.make_config(50, shock_fused=1)
    probs = [batch.problem(50, ((4.4, 6e8), (1.4, 0.0))), batch.problem(60, ((4.4, 6e8), (1.4, 0.0)))]
    with pytest.raises(spb6.SPB6Error):
        batch.Batch(probs, scheme)
