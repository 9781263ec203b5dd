"""The CPU oracle against the reference's own archived outputs (tests/golden/ref_profiles.npz, extracted from the
reference's data/ folder by tests/golden/make_ref_profiles.py).

In its PC mode (src/spb6_physics.h closure path, B = 4096 block sums, libm transcendental functions) the oracle must
reproduce every archived profile bit for bit: these 1D runs exercise the explicit MUSCL-Rusanov stage, the stage
preparation, the shock budget (strong and weak branches), the p closure, Newton with the finite-difference Jacobian,
Thomas, the line search, the start-up ramp, rejects (case 83) and the LSDIRK2 driver.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_profiles.npz")
COLS = ["p", "u", "v", "rho", "alpha2", "Y2", "rho1", "rho2", "p1", "p2"]
DIAG = {"p": 0, "u": 1, "v": 2, "rho": 3, "alpha2": 4, "Y2": 5, "rho1": 6, "rho2": 7, "p1": 12, "p2": 13}
KEYS = ["c80_200", "c81_200", "c82_200", "c83_200", "c84_200", "c85_200", "c86_200", "c87_200", "c88_200", "c5_100", "c4_200"]
# archived run statistics (data/*/note.md): steps, first-order steps, rejects
NOTES = {"c81_200": (17, 17, 0), "c83_200": (317, 37, 8), "c84_200": (150, 0, 0), "c85_200": (18, 18, 0),
         "c86_200": (287, 42, 0), "c87_200": (120, 6, 0), "c88_200": (24, 11, 0), "c80_200": (537, 42, 0)}


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


@pytest.mark.parametrize("key", KEYS)
def test_oracle_reproduces_archived_profile(oracle_lib, gold, key):
    O = oracle_lib
    case, n = (int(v) for v in gold[key + "/meta"])
    cfg, U, tend = O.case_init(case, n)
    O.reference_mode(cfg, "PC")
    o = O.Oracle(cfg)
    o.set_state(U)
    assert o.run(tend) == 0
    D = O.cell_diag(cfg, o.state())
    for name in COLS:
        a, b = D[:, DIAG[name]], gold[f"{key}/{name}"]
        assert np.array_equal(a, b), f"{key}: column {name} differs in {np.count_nonzero(a != b)} cells"
    st = o.stats()
    if key in NOTES:
        steps, rk1, rej = NOTES[key]
        # the note's step count includes the final partial step the same way for all cases but c83 (one off, archived
        # from a different invocation); first-order steps and rejects match exactly
        assert abs(st["steps"] - steps) <= 1 and st["rk1"] == rk1 and st["rejects"] == rej


def test_pc_omp_path_spread_is_reported(oracle_lib, gold):
    """The PC_omp iteration paths (closure Newton from the guess, weak-shock-centred share bracket) change where each
    root stops inside the reference's own stopping tolerance; on the water-air shock tube the whole-run spread stays
    far below the discretisation error (archived case 87 N=200)."""
    O = oracle_lib
    cfg, U, tend = O.case_init(87, 200)
    O.reference_mode(cfg, "PC_omp")
    o = O.Oracle(cfg)
    o.set_state(U)
    assert o.run(tend) == 0
    p = O.cell_diag(cfg, o.state())[:, 0]
    ref = gold["c87_200/p"]
    l1 = np.abs(p - ref).sum() / np.abs(ref).sum()
    assert l1 < 1e-4
