"""The exponential relaxation form (RELAX_FORM 1, PHY:39-41) of the theta closure (PHY:685: T2 - T1 decays by
exp(-tau theta S)) and the nu closure (PHY:721-745: Y moves from the transported fraction towards Y_eq along the secant
of g2 - g1). No archived run uses it; it is pinned by its portable e^x - 1 against libm and by the limits both forms
share: a vanishing rate gives the mechanical closure's run, an infinite one the instantaneous closure's, and between
them the two forms differ."""
import math

import numpy as np
import pytest


def test_portable_expm1_against_libm(oracle_lib):
    L = oracle_lib.lib()
    xs = np.concatenate([np.linspace(-0.5, 0.5, 4001), 10 ** np.linspace(-300, -0.31, 600), -(10 ** np.linspace(-300, -0.31, 600)),
                         np.linspace(-40, 40, 801)])
    for x in xs:
        a, b = L.ora_pexpm1(float(x)), math.expm1(float(x))
        assert abs(a - b) <= 2 * abs(b) * 2 ** -52, (x, a, b)
    assert L.ora_pexpm1(-math.inf) == -1.0


def _run(O, relax, steps=50, **kw):
    cfg, U, _ = O.case_init(30, 400)
    cfg.relax = relax
    for k, v in kw.items():
        setattr(cfg, k, v)
    o = O.Oracle(cfg)
    o.set_state(U)
    assert o.run(1.0, steps) == 0
    return o.state()


@pytest.mark.parametrize("form", [0, 1])
def test_theta_limits(oracle_lib, form):
    """level 1 on Pelanti's NASG tube (400 cells, 50 steps): theta = 1e-12 gives the level-0 run (the closure's heat
    exchange vanishes) and theta = 1e15 the instantaneous pT run, in both forms"""
    O = oracle_lib
    U0 = _run(O, 0)
    Uinf = _run(O, 1)
    Us = _run(O, 1, theta_rate=1e-12, relax_form=form)
    Ub = _run(O, 1, theta_rate=1e15, relax_form=form)
    assert np.abs(Us - U0).max() <= 1e-6 * np.abs(U0).max()
    assert np.abs(Ub - Uinf).max() <= 1e-6 * np.abs(Uinf).max()


def test_forms_differ_at_finite_rates(oracle_lib):
    O = oracle_lib
    a = _run(O, 1, theta_rate=2000.0, relax_form=0)
    b = _run(O, 1, theta_rate=2000.0, relax_form=1)
    assert np.abs(a - b).max() > 0
    c = _run(O, 2, nu_rate=1e-4, relax_form=0)
    d = _run(O, 2, nu_rate=1e-4, relax_form=1)
    assert np.abs(c - d).max() > 0
