"""Relaxation levels 1-3 on the GPU: whole runs bit-identical to the oracle in its device mode (state, pressure and
counters; the pT and pTg closures' root finders and the saturation line use the portable exp / log shared with the
device; the saturation table is built on the host with them), and the reference's archived profiles of the
Pelanti-Shyue water cavitation tube (5000 cells, t = 3.2 ms) reproduced: level 1 to round-off, levels 2-3 within the
spread the oracle measures live for the reference's own variants (its OpenMP program's paths, and its serial paths with
the portable exp and log): the phase-change decisions flip on last-ulp changes (test_oracle_relax)."""
import os

import numpy as np
import pytest

from tests.helpers import bits_diff, oracle_cfg_from, config_from_oracle, field_l1

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _pair_run(scfg, U, steps, activate=False):
    from diffusionimex_amd import spb6
    from oracle import oracle as O
    o = O.Oracle(oracle_cfg_from(scfg, O))
    o.set_state(U)
    assert o.run(1.0, steps) == 0
    st = spb6.Stepper(scfg)
    st.set_state(U)
    assert st.run(1.0, steps) == 0
    if activate:
        o.L.ora_relax_activate(o.h)
        assert st.relax_activate() == 0
    return o, st


def _check(o, st, name):
    so, ss = o.stats(), st.stats()
    for k in ("steps", "rk1", "rejects", "stages", "newton_it", "krylov_it", "floor_acc", "linsolves"):
        assert so[k] == ss[k], f"{name}: {k} {so[k]} vs {ss[k]}"
    Uo, Us = o.state(), st.state()
    assert bits_diff(Us, Uo) == 0, f"{name}: state differs in {bits_diff(Us, Uo)} values; per-field L1 {field_l1(Us, Uo)}"
    assert bits_diff(st.pressure()[0], o.pressure()[0]) == 0


def test_level1_tube_bitwise():
    from oracle import oracle as O
    ocfg, U, _ = O.case_init(1, 800)
    scfg = config_from_oracle(ocfg, relax=1)
    o, st = _pair_run(scfg, U, 60, activate=True)
    _check(o, st, "tube800")


def test_level1_2d_bitwise():
    """A 2D water-vapour field: the P-S liquid / vapour EOS, a vapour pocket with a radial expansion, walls in y."""
    from diffusionimex_amd import cases, spb6
    from oracle import oracle as O
    nx, ny = 40, 28
    ocfg, _, _ = O.case_init(1, nx, ny)
    scfg = config_from_oracle(ocfg, relax=1)
    scfg.bc[2] = scfg.bc[3] = spb6.BC_WALL
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)
    r2 = (X - 0.5) ** 2 + (Y - 0.5) ** 2
    a1 = np.where(r2 < 0.04, 0.9, 1 - 1e-2)
    ph = [[ocfg.ph[k][q] for q in range(6)] for k in range(2)]
    T0 = 354.728
    v = lambda k, p, T: (ph[k][0] - 1) * ph[k][4] * T / (p + ph[k][1]) + ph[k][5]
    U = cases.prim2cell(ph, a1, 1.0 / v(0, 1e5, T0), 1.0 / v(1, 1e5, T0), 3.0 * (X - 0.5), 3.0 * (Y - 0.5), 1e5, 1e5)
    o, st = _pair_run(scfg, U, 25, activate=True)
    _check(o, st, "pocket40x28")


def test_level1_archived_profile():
    """The reference's archived p-pT profile (tests/golden/ref_profiles.npz "c1_5000_r1"). The oracle reproduces it bit for
    bit with libm's exp and log in both the serial and the OpenMP program's paths (test_oracle_relax; checked here by
    hand for the PC_omp paths); the device, which equals the oracle's device mode bit for bit, uses the portable exp and
    log shared with the oracle, and the pT closure's bracket search moves between p and y = ln(p - plo) through them, so
    each root stops a few ulps away and the 188 steps end at 3.4e-12 relative L1 in p (3.5e-12 in p1, 2.7e-14 in alpha2,
    1.6e-14 in u, 7e-16 in rho; the oracle's device mode gives the same). Held to 1e-11 per column, with the archived
    step, first-order-step and Newton counts exactly."""
    from diffusionimex_amd import spb6
    from oracle import oracle as O
    gold = np.load(os.path.join(HERE, "golden", "ref_profiles.npz"))
    ocfg, U, tend = O.case_init(1, 5000)
    st = spb6.Stepper(config_from_oracle(ocfg, relax=1))
    st.set_state(U)
    assert st.run(tend) == 0
    assert st.relax_activate() == 0
    s = st.stats()
    assert (s["steps"], s["rk1"], s["newton_it"]) == (188, 33, 695)
    D = O.cell_diag(ocfg, st.state())
    for name, col in {"p": 0, "u": 1, "rho": 3, "alpha2": 4, "p1": 12, "p2": 13}.items():
        a, b = D[:, col], gold[f"c1_5000_r1/{name}"]
        assert np.abs(a - b).sum() / np.abs(b).sum() <= 1e-11, name


@pytest.mark.parametrize("relax", [2, 3])
def test_levels23_tube_bitwise(relax):
    """the whole 5000-cell tube (188 steps; vapour forms in the rarefaction) at level 2 (p-pT-pTg) and 3 (p-pTg)"""
    from oracle import oracle as O
    from diffusionimex_amd import spb6
    ocfg, U, tend = O.case_init(1, 5000)
    scfg = config_from_oracle(ocfg, relax=relax)
    o = O.Oracle(oracle_cfg_from(scfg, O))
    o.set_state(U)
    assert o.run(tend) == 0
    st = spb6.Stepper(scfg)
    st.set_state(U)
    assert st.run(tend) == 0
    o.L.ora_relax_activate(o.h)
    assert st.relax_activate() == 0
    assert st.stats()["steps"] == 188
    _check(o, st, f"tube5000_r{relax}")


@pytest.mark.parametrize("relax", [2, 3])
def test_levels23_2d_bitwise(relax):
    """the 2D vapour pocket of the level-1 test under expansion, at levels 2 and 3"""
    from diffusionimex_amd import cases, spb6
    from oracle import oracle as O
    nx, ny = 40, 28
    ocfg, _, _ = O.case_init(1, nx, ny)
    scfg = config_from_oracle(ocfg, relax=relax)
    scfg.bc[2] = scfg.bc[3] = spb6.BC_WALL
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)
    r2 = (X - 0.5) ** 2 + (Y - 0.5) ** 2
    a1 = np.where(r2 < 0.04, 0.9, 1 - 1e-2)
    ph = [[ocfg.ph[k][q] for q in range(6)] for k in range(2)]
    T0 = 354.728
    v = lambda k, p, T: (ph[k][0] - 1) * ph[k][4] * T / (p + ph[k][1]) + ph[k][5]
    U = cases.prim2cell(ph, a1, 1.0 / v(0, 1e5, T0), 1.0 / v(1, 1e5, T0), 30.0 * (X - 0.5), 30.0 * (Y - 0.5), 1e5, 1e5)
    o, st = _pair_run(scfg, U, 25, activate=True)
    _check(o, st, f"pocket40x28_r{relax}")


@pytest.mark.parametrize("relax", [2, 3])
def test_levels23_archived_profiles(relax):
    """The archived p-pT-pTg and p-pTg profiles: per column within twice the larger of the reference's own variants'
    distances to the archive (the oracle in the OpenMP program's paths with libm, and in the serial program's paths with
    the portable exp and log), measured here on the CPU, plus 1e-12; the steps and first-order steps exactly."""
    from diffusionimex_amd import spb6
    from oracle import oracle as O
    gold = np.load(os.path.join(HERE, "golden", "ref_profiles.npz"))
    key = f"c1_5000_r{relax}"
    cols = {"p": 0, "u": 1, "rho": 3, "alpha2": 4, "Y2": 5, "p1": 12, "p2": 13}
    ocfg, U, tend = O.case_init(1, 5000)
    st = spb6.Stepper(config_from_oracle(ocfg, relax=relax))
    st.set_state(U)
    assert st.run(tend) == 0
    assert st.relax_activate() == 0
    s = st.stats()
    assert (s["steps"], s["rk1"]) == (188, 33)
    D = O.cell_diag(ocfg, st.state())

    def spread(setup):
        cfg, U0, t_end = O.case_init(1, 5000)
        setup(cfg)
        cfg.relax = relax
        o = O.Oracle(cfg)
        o.set_state(U0)
        assert o.run(t_end) == 0
        o.L.ora_relax_activate(o.h)
        Do = O.cell_diag(cfg, o.state())
        return {k: np.abs(Do[:, c] - gold[f"{key}/{k}"]).sum() / np.abs(gold[f"{key}/{k}"]).sum() for k, c in cols.items()}

    omp = spread(lambda c: O.reference_mode(c, "PC_omp"))
    port = spread(lambda c: (O.reference_mode(c, "PC"), setattr(c, "portable_math", 1)))
    for k, c in cols.items():
        got = np.abs(D[:, c] - gold[f"{key}/{k}"]).sum() / np.abs(gold[f"{key}/{k}"]).sum()
        assert got <= 2.0 * max(omp[k], port[k]) + 1e-12, (relax, k, got, omp[k], port[k])


@pytest.mark.parametrize("relax", [0, 1, 2])
def test_nasg_tube_bitwise(relax):
    """Pelanti's water cavitation tube (case 30: NASG liquid with covolume, vapour; 5000 cells, 178 steps) at levels
    0-2, bit-identical to the oracle's device mode"""
    from diffusionimex_amd import spb6
    from oracle import oracle as O
    ocfg, U, tend = O.case_init(30, 5000)
    scfg = config_from_oracle(ocfg, relax=relax)
    o = O.Oracle(oracle_cfg_from(scfg, O))
    o.set_state(U)
    assert o.run(tend) == 0
    st = spb6.Stepper(scfg)
    st.set_state(U)
    assert st.run(tend) == 0
    if relax:
        o.L.ora_relax_activate(o.h)
        assert st.relax_activate() == 0
    assert st.stats()["steps"] == 178
    _check(o, st, f"nasg5000_r{relax}")


@pytest.mark.parametrize("rates", [(1, 2000.0, 0.0), (2, 0.0, 1e-4)], ids=["theta2000", "nu1e-4"])
def test_finite_rates_bitwise(rates):
    """the finite-rate closures on Pelanti's tube (case 30): level 1 with theta = 2000 and level 2 with nu = 1e-4 (the
    runs the reference archived), bit-identical to the oracle's device mode"""
    from diffusionimex_amd import spb6
    from oracle import oracle as O
    relax, th, nu = rates
    ocfg, U, tend = O.case_init(30, 5000)
    scfg = config_from_oracle(ocfg, relax=relax, theta_rate=th, nu_rate=nu)
    oc = oracle_cfg_from(scfg, O)
    assert (oc.theta_rate, oc.nu_rate) == (th if th > 0 else -1.0, nu if nu > 0 else -1.0)
    o = O.Oracle(oc)
    o.set_state(U)
    assert o.run(tend) == 0
    st = spb6.Stepper(scfg)
    st.set_state(U)
    assert st.run(tend) == 0
    o.L.ora_relax_activate(o.h)
    assert st.relax_activate() == 0
    _check(o, st, f"nasg5000_r{relax}_th{th}_nu{nu}")


@pytest.mark.parametrize("case,relax", [(3, 0), (31, 0), (31, 2)], ids=["dodecane_sg_r0", "dodecane_nasg_r0", "dodecane_nasg_r2"])
def test_dodecane_tube_bitwise(case, relax):
    """the dodecane liquid / vapour shock tubes (2000 cells, t = 473 us): the SG form (case 3, material CFL 0.05, 2675
    steps) whose strong shock into the vapour runs the profile shooting in every shocked cell, and the NASG form (case 31,
    covolume 7.51e-4) at levels 0 and 2, bit-identical to the oracle's device mode"""
    from diffusionimex_amd import spb6
    from oracle import oracle as O
    ocfg, U, tend = O.case_init(case, 2000)
    scfg = config_from_oracle(ocfg, relax=relax)
    o = O.Oracle(oracle_cfg_from(scfg, O))
    o.set_state(U)
    assert o.run(tend) == 0
    st = spb6.Stepper(scfg)
    st.set_state(U)
    assert st.run(tend) == 0
    if relax:
        o.L.ora_relax_activate(o.h)
        assert st.relax_activate() == 0
    _check(o, st, f"dodecane{case}_r{relax}")


@pytest.mark.parametrize("relax,mu,steps", [(0, 1e3, 60), (0, 1e8, 40), (1, 1e3, 40)], ids=["r0_mu1e3", "r0_mu1e8", "r1_mu1e3"])
def test_finite_mu_bitwise(relax, mu, steps):
    """the finite mechanical rate (the mu closure with the stage's frozen impedances) on the 2 m/s cavitation tube (800
    cells): at level 0 every cell is of the mechanical type CL_MU; at level 1 the cells outside eps_I <= alpha_1 <=
    1 - eps_I are. With mu = 1e3 the run takes rejected attempts (trial transfers that leave the vapour without pressure)
    as the oracle does; bit-identical to the oracle's device mode, counters included"""
    from oracle import oracle as O
    ocfg, U, _ = O.case_init(1, 800)
    scfg = config_from_oracle(ocfg, relax=relax, mu_rate=mu)
    oc = oracle_cfg_from(scfg, O)
    assert oc.mu_rate == mu
    o, st = _pair_run(scfg, U, steps, activate=relax > 0)
    _check(o, st, f"tube800_r{relax}_mu{mu}")


def test_finite_mu_2d_bitwise():
    """a 2D vapour pocket under expansion (40 x 28, walls in y) at level 0 with mu = 1e4: the general path's 2D stage
    with CL_MU in every cell"""
    from diffusionimex_amd import cases, spb6
    from oracle import oracle as O
    nx, ny = 40, 28
    ocfg, _, _ = O.case_init(1, nx, ny)
    scfg = config_from_oracle(ocfg, mu_rate=1e4)
    scfg.bc[2] = scfg.bc[3] = spb6.BC_WALL
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)
    a1 = np.where((X - 0.5) ** 2 + (Y - 0.5) ** 2 < 0.04, 0.9, 1 - 1e-2)
    ph = [[ocfg.ph[k][q] for q in range(6)] for k in range(2)]
    T0 = 354.728
    v = lambda k, p, T: (ph[k][0] - 1) * ph[k][4] * T / (p + ph[k][1]) + ph[k][5]
    U = cases.prim2cell(ph, a1, 1.0 / v(0, 1e5, T0), 1.0 / v(1, 1e5, T0), 3.0 * (X - 0.5), 3.0 * (Y - 0.5), 1e5, 1.2e5)
    o, st = _pair_run(scfg, U, 20)
    _check(o, st, "pocket40x28_mu1e4")


@pytest.mark.parametrize("relax", [2, 3])
def test_nu_law_bitwise(relax):
    """Pelanti's state-dependent transfer rate (the nu law, psat from the device's saturation table) on the NASG water
    tube (case 30, 1000 cells, 60 steps) at levels 2 and 3: each attempt's activation freezes the chemical cells' rates
    from their pT equilibria; bit-identical to the oracle's device mode"""
    from oracle import oracle as O
    ocfg, U, _ = O.case_init(30, 1000)
    scfg = config_from_oracle(ocfg, relax=relax, nu_law=1)
    o, st = _pair_run(scfg, U, 60, activate=True)
    _check(o, st, f"nasg1000_r{relax}_nulaw")


@pytest.mark.parametrize("rates", [(1, 2000.0, 0.0), (2, 0.0, 1e-4)], ids=["theta2000", "nu1e-4"])
def test_exponential_form_bitwise(rates):
    """the exponential form of the theta and nu closures (RELAX_FORM 1) on Pelanti's NASG tube (5000 cells, to t_end),
    the portable e^x - 1 and e^x shared with the oracle: bit-identical to the oracle's device mode"""
    from diffusionimex_amd import spb6
    from oracle import oracle as O
    relax, th, nu = rates
    ocfg, U, tend = O.case_init(30, 5000)
    scfg = config_from_oracle(ocfg, relax=relax, theta_rate=th, nu_rate=nu, relax_form=1)
    o = O.Oracle(oracle_cfg_from(scfg, O))
    o.set_state(U)
    assert o.run(tend) == 0
    st = spb6.Stepper(scfg)
    st.set_state(U)
    assert st.run(tend) == 0
    o.L.ora_relax_activate(o.h)
    assert st.relax_activate() == 0
    _check(o, st, f"nasg5000_form1_r{relax}")


def test_mu_exponential_form_bitwise():
    """the mu closure's exponential form (the equilibrium transfer, then R = R_eq (1 - exp(-tau mu D))) at level 0 with
    mu = 1e3 on the 2 m/s cavitation tube (800 cells, 60 steps): bit-identical to the oracle's device mode"""
    from oracle import oracle as O
    ocfg, U, _ = O.case_init(1, 800)
    scfg = config_from_oracle(ocfg, mu_rate=1e3, relax_form=1)
    o, st = _pair_run(scfg, U, 60)
    _check(o, st, "tube800_mu1e3_form1")
