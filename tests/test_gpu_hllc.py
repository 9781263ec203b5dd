"""The explicit HLLC scheme (src/SPB6_explicit.cpp) on the GPU against the oracle, bit for bit.

The oracle runs in its device mode (the PC_omp closure path, portable exp/log/pow), the mode the HIP kernels share; in
the reference's own mode it reproduces the reference's published results for the scheme (test_oracle_hllc). Whole runs
of the time loop are compared: state, steps, rejects and the relaxation's closure failures.
"""
import numpy as np
import pytest

from tests.helpers import bits_diff, oracle_cfg_from, field_l1

pytestmark = pytest.mark.gpu


def _cases():
    from diffusionimex_amd import cases
    from diffusionimex_amd.spb6 import BC_PERIODIC
    out = []
    cfg, U = cases.water_air_tube(200)
    out.append(("tube200", cfg, U, 150))
    cfg, U = cases.shock_bubble(48, 24, seed=3)
    out.append(("bubble48x24", cfg, U, 40))
    cfg, U = cases.perturbed_interface(37, 29, seed=2)   # odd sizes: partial blocks in both directions
    out.append(("interface37x29", cfg, U, 30))
    cfg, U = cases.cavitation_2d(33, 21, seed=4, bc=(BC_PERIODIC,) * 4)
    out.append(("cavitation33x21_periodic", cfg, U, 30))
    cfg, U = cases.shock_bubble(40, 24, seed=3, order=1)   # first order: the face states are the cell primitives
    out.append(("bubble40x24_first_order", cfg, U, 30))
    return out


@pytest.mark.parametrize("idx", [0, 1, 2, 3, 4])
def test_hllc_run_bitwise(idx):
    from diffusionimex_amd import spb6
    from oracle import oracle as O
    name, cfg, U, steps = _cases()[idx]
    o = O.Oracle(oracle_cfg_from(cfg, O))
    o.set_state(U)
    f0 = o.L.ora_hllc_closure_fails()
    rco = o.hllc_run(1.0, steps, 0.4)
    st = spb6.Stepper(cfg)
    st.set_state(U)
    rcs = st.hllc_run(1.0, steps, 0.4)
    assert rco == 0 and rcs == 0, (name, rco, rcs)
    so, ss = o.stats(), st.stats()
    for k in ("steps", "rejects"):
        assert so[k] == ss[k], f"{name}: {k} {so[k]} vs {ss[k]}"
    assert ss["closure_fails"] == o.L.ora_hllc_closure_fails() - f0
    assert so["t"] == ss["t"]
    Uo, Us = o.state(), st.state()
    assert bits_diff(Us, Uo) == 0, f"{name}: state differs in {bits_diff(Us, Uo)} values; per-field L1 {field_l1(Us, Uo)}"


def test_hllc_rejects_bitwise():
    """A step far beyond the acoustic limit fails and is halved until it passes, as on the oracle."""
    from diffusionimex_amd import cases, spb6
    from oracle import oracle as O
    cfg, U = cases.water_air_tube(200)
    o = O.Oracle(oracle_cfg_from(cfg, O))
    o.set_state(U)
    assert o.hllc_run(1.0, 12, 3.0) == 0
    st = spb6.Stepper(cfg)
    st.set_state(U)
    assert st.hllc_run(1.0, 12, 3.0) == 0
    so, ss = o.stats(), st.stats()
    assert so["rejects"] > 0 and so["rejects"] == ss["rejects"] and so["steps"] == ss["steps"]
    assert bits_diff(st.state(), o.state()) == 0


def test_hllc_water_air_tube_published_results_on_gpu():
    """docs/results.md figure 8: max u / u* = 1.002, shock at x = 0.9385 (1000 cells, 240 us)."""
    from diffusionimex_amd import spb6
    from oracle import oracle as O
    ocfg, U, tend = O.case_init(5, 1000)
    from tests.helpers import config_from_oracle
    st = spb6.Stepper(config_from_oracle(ocfg))
    st.set_state(U)
    assert st.hllc_run(tend, -1, 0.4) == 0
    D = O.cell_diag(ocfg, st.state().reshape(-1, 8))
    u = D[:, 1]
    x = (np.arange(1000) + 0.5) / 1000
    assert round(u.max() / 492.1, 3) == 1.002
    assert abs(x[np.where(u > 0.5 * 492.1)[0].max()] + 0.5e-3 - 0.9385) <= 1e-3
