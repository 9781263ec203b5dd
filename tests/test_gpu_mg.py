"""Mie-Grueneisen phases on the GPU (libspb6_hip_mg.so): whole runs bit-identical to the oracle in its device mode
(the PHO iteration paths of the MG closure, the portable exp and pow in the reference curves): the nitromethane tube
(Cochran-Chan, case 60, which the oracle reproduces from the reference's archive bit for bit), the detonation gases
(JWL) | water (NASG) tube (case 61), and a 2D JWL bubble in water between walls."""
import numpy as np
import pytest

from tests.helpers import bits_diff, config_from_oracle, field_l1, oracle_cfg_from

pytestmark = pytest.mark.gpu


def _check(o, st, name):
    so, ss = o.stats(), st.stats()
    for k in ("steps", "rk1", "rejects", "stages", "newton_it", "krylov_it", "floor_acc", "linsolves"):
        assert so[k] == ss[k], f"{name}: {k} {so[k]} vs {ss[k]}"
    Uo, Us = o.state(), st.state()
    assert bits_diff(Us, Uo) == 0, f"{name}: state differs in {bits_diff(Us, Uo)} values; per-field L1 {field_l1(Us, Uo)}"
    assert bits_diff(st.pressure()[0], o.pressure()[0]) == 0


def _pair(scfg, U, tend, steps=-1):
    from diffusionimex_amd import spb6
    from oracle import oracle as O
    o = O.Oracle(oracle_cfg_from(scfg, O))
    o.set_state(U)
    assert o.run(tend, steps) == 0
    st = spb6.Stepper(scfg)
    st.set_state(U)
    assert st.run(tend, steps) == 0
    return o, st


@pytest.mark.parametrize("case", [60, 61], ids=["nitromethane_cochran_chan", "jwl_water"])
def test_mg_tube_bitwise(case):
    from oracle import oracle as O
    ocfg, U, tend = O.case_init(case, 500)
    scfg = config_from_oracle(ocfg)
    o, st = _pair(scfg, U, tend)
    _check(o, st, f"case{case}")


def test_mg_2d_bubble_bitwise():
    """a JWL bubble (4.6406e10 Pa) in NASG water at 1e5 Pa, 48 x 32 cells between walls in y, 25 steps: the 2D explicit
    stage, the cross-stencil shock budget with the MG Hugoniot and the 2D BiCGStab on MG states"""
    from diffusionimex_amd import cases, spb6
    from oracle import oracle as O
    nx, ny = 48, 32
    ocfg, _, _ = O.case_init(61, nx, ny)
    ocfg.x0, ocfg.x1, ocfg.y0, ocfg.y1 = -1.0, 1.0, -0.6, 0.6
    scfg = config_from_oracle(ocfg)
    scfg.bc[2] = scfg.bc[3] = spb6.BC_WALL
    x = -1.0 + (np.arange(nx) + 0.5) * 2.0 / nx
    y = -0.6 + (np.arange(ny) + 0.5) * 1.2 / ny
    X, Y = np.meshgrid(x, y)
    inb = (X + 0.3) ** 2 + Y ** 2 < 0.3 ** 2
    U = np.empty((nx * ny, 8))
    cfg1, U1, _ = O.case_init(61, 2)   # cell 0: the gas state, cell 1: the water state of the tube
    U[:] = np.where(inb.reshape(-1, 1), U1[0], U1[1])
    o, st = _pair(scfg, U, 1.0, 25)
    _check(o, st, "bubble48x32")
