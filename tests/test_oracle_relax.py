"""Relaxation level 1 (p-pT: thermal equilibrium in the mixture cells, the pT closure PHY:660-672 with its root finders
PHY:183-225, and the step's RelaxationActivation PC:537-560) in the oracle, against the reference's archived
full-precision profiles of the Pelanti-Shyue water cavitation tubes (data/PS2014_water_cavitation_tube_2mps and _500mps,
5000 cells; extracted into tests/golden/ref_profiles.npz). In its PC mode the oracle reproduces them bit for bit, with
the archived step, first-order-step and Newton counts (data/*/note.md)."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_profiles.npz")
DIAG = {"p": 0, "u": 1, "rho": 3, "alpha2": 4, "p1": 12, "p2": 13}
NOTES = {"c1_5000_r0": (188, 33, 719), "c1_5000_r1": (188, 33, 695), "c2_5000_r1": (7252, 4, 14504)}


def _check(O, gold, key):
    case, n, relax = (int(v) for v in gold[key + "/meta"])
    cfg, U, tend = O.case_init(case, n)
    O.reference_mode(cfg, "PC")
    cfg.relax = relax
    o = O.Oracle(cfg)
    o.set_state(U)
    assert o.run(tend) == 0
    o.L.ora_relax_activate(o.h)   # the reference's main activates once more before its output (PC:1300)
    D = O.cell_diag(cfg, o.state())
    for name, col in DIAG.items():
        a, b = D[:, col], gold[f"{key}/{name}"]
        assert np.array_equal(a, b), f"{key}: column {name} differs in {np.count_nonzero(a != b)} cells"
    st = o.stats()
    assert (st["steps"], st["rk1"], st["newton_it"]) == NOTES[key]


@pytest.mark.parametrize("key", ["c1_5000_r0", "c1_5000_r1"])
def test_cavitation_tube_reproduced_bit_for_bit(oracle_lib, key):
    _check(oracle_lib, np.load(GOLD), key)


@pytest.mark.slow
@pytest.mark.skipif(not os.environ.get("SPB6_SLOW"), reason="7252 steps of 5000 cells (about a minute); SPB6_SLOW=1 runs it")
def test_fast_cavitation_tube_reproduced_bit_for_bit(oracle_lib):
    _check(oracle_lib, np.load(GOLD), "c2_5000_r1")


def test_level1_types_and_projection(oracle_lib):
    """After an activation, every mixture cell of the tube is in thermal equilibrium: both phases at one temperature
    (to round-off) and one pressure, the pure-ish cells (alpha_1 outside [eps_I, 1 - eps_I]) untouched."""
    O = oracle_lib
    cfg, U, tend = O.case_init(1, 400)
    cfg.relax = 1
    o = O.Oracle(cfg)
    o.set_state(U)
    assert o.run(tend, 20) == 0
    o.L.ora_relax_activate(o.h)
    D = O.cell_diag(cfg, o.state())
    T1, T2, p1, p2 = D[:, 10], D[:, 11], D[:, 12], D[:, 13]
    assert np.max(np.abs(T1 - T2) / T1) < 1e-12
    assert np.max(np.abs(p1 - p2) / (np.abs(p1) + 1e9)) < 1e-12
