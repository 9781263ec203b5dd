"""The state-dependent mass-transfer rate of Pelanti 2022 eq. (84) in the oracle (NU_LAW, PHY:34-36, PHY:162-177,
PC:539-556, PC:671): nu = c_r alpha_v^0.6 ((psat(T) - p) / (p_c - psat(T)))^1.76 where the liquid is superheated
(p < psat(T)), frozen per cell at the step's activation from the cell's pT equilibrium, in the chemical cells' nu closure.
No archived run of the reference covers it in a form this build reads (the Bartak tube's archive is a probe series with a
pressure outlet), so it is pinned by what fixes it: psat inverts the saturation line, the rate vanishes where the law
says it does and scales with its exponents, and a run's chemical cells carry the rates the law gives for their
equilibria."""
import ctypes as C

import numpy as np
import pytest


def _ctx(O, relax=2, **kw):
    cfg, U, tend = O.case_init(30, 400)
    O.reference_mode(cfg, "PC")
    cfg.relax, cfg.nu_law = relax, 1
    for k, v in kw.items():
        setattr(cfg, k, v)
    o = O.Oracle(cfg)
    o.set_state(U)
    return o, cfg


def test_psat_inverts_the_saturation_line(oracle_lib):
    O = oracle_lib
    o, _ = _ctx(O)
    L = O.lib()
    for T in np.linspace(300.0, 600.0, 13):
        ps = L.ora_psat(o.h, T)
        assert 0 < ps < 2.2e7
        assert abs(L.ora_tsat(o.h, ps) - T) <= 1e-9 * T, (T, ps)
    assert L.ora_psat(o.h, 1.0) == 0.0 and np.isinf(L.ora_psat(o.h, 5000.0))


def test_nu_law_zeros_and_exponents(oracle_lib):
    O = oracle_lib
    o, _ = _ctx(O)
    L = O.lib()
    T = 400.0
    ps = L.ora_psat(o.h, T)
    assert L.ora_nu_law(o.h, ps * 1.01, T, 0.1) == 0.0          # not superheated
    assert L.ora_nu_law(o.h, ps * 0.5, T, 0.0) == 0.0           # no vapour
    n1 = L.ora_nu_law(o.h, ps * 0.5, T, 0.1)
    n2 = L.ora_nu_law(o.h, ps * 0.5, T, 0.2)
    assert n1 > 0 and abs(n2 / n1 - 2.0 ** 0.6) <= 1e-13
    d1, d2 = ps - ps * 0.5, ps - ps * 0.75
    n3 = L.ora_nu_law(o.h, ps * 0.75, T, 0.1)
    assert abs(n1 / n3 - (d1 / d2) ** 1.76) <= 1e-12
    o2, _ = _ctx(O, nu_cr=3.0)
    assert abs(O.lib().ora_nu_law(o2.h, ps * 0.5, T, 0.1) / n1 - 3.0) <= 1e-15


def test_run_cells_carry_the_law(oracle_lib):
    """Pelanti's NASG water tube (case 30, 400 cells) at level 2 with the law: 40 steps run, the chemical cells of the
    last activation carry positive rates where the rarefaction superheats the liquid, and the state differs from the
    instantaneous-rate run"""
    O = oracle_lib
    o, cfg = _ctx(O)
    assert o.run(1.0, 40) == 0
    o.L.ora_relax_activate(o.h)
    nu = np.empty(400)
    assert O.lib().ora_cell_nu(o.h, nu.ctypes.data_as(C.POINTER(C.c_double))) == 400
    assert (nu > 0).sum() > 0 and np.all(nu >= 0)
    c2, U, _ = O.case_init(30, 400)
    O.reference_mode(c2, "PC")
    c2.relax = 2
    oi = O.Oracle(c2)
    oi.set_state(U)
    assert oi.run(1.0, 40) == 0
    assert np.abs(o.state() - oi.state()).max() > 0
