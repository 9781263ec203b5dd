"""The fused form of the strong-shock profile integration (shock_fused = 1: prof_rhs_fused / prof_integrate_fused in the
oracle, prof_step_fused in the HIP kernel) against the reference's arithmetic of the same ODE (PHY:302-341), on the
CPU.

Pins:
- the right-hand side: on random profile states of four phase pairs (water-air, air-helium, a covolume (NASG) liquid
  with air, two stiff liquids) both forms are compared with the reference's formula evaluated in extended precision
  (np.longdouble, 64-bit mantissa). The fused form must be as accurate as the reference's own rounding (error at most
  4x the reference form's plus 16 ulp of the terms' scale), so a dropped or transposed term (e1 x1 for e1 x0, den x1 for
  den x0, the sign of n0 num) fails by orders of magnitude;
- the shares: shock_share in both forms on random strong shocks: the same outcome, lam within 1e-9 and q within 1e-9
  relative (the shooting stops at a miss below 1e-10; a share near 0 or 1 moves the logit t, not lam, at that level);
- whole runs: the water-air tube (case 5, 200 cells, to its end) and a 2D shock-bubble (12 steps) in both forms: the
  same step and stage counts, and per field the fused run no farther from the reference arithmetic's run than 10x
  what a one-ulp change of one cell's initial energy moves that run (plus 1e-12 relative L1). Newton's tolerance-level
  solves carry such a change to 1e-6 relative in two steps of the bubble, so no fixed bound in between would hold.
"""
import ctypes as C

import numpy as np
import pytest

LD = np.longdouble
PAIRS = {
    "water_air": ([4.4, 6e8, 0, 0, 1000.0, 0], [1.4, 0.0, 0, 0, 718.0, 0], 1e5, 1e9, 1000.0, 1.2),
    "air_helium": ([1.4, 0.0, 0, 0, 718.0, 0], [1.67, 0.0, 0, 0, 3100.0, 0], 1e5, 4e5, 1.2, 0.17),
    "nasg_air": ([1.19, 6.2e8, 0, 0, 3610.0, 6.7e-4], [1.4, 0.0, 0, 0, 718.0, 0], 1e5, 5e8, 960.0, 1.2),
    "stiff_pair": ([4.4, 6e8, 0, 0, 1000.0, 0], [2.8, 8.5e8, 0, 0, 1500.0, 0], 1e5, 2e9, 1000.0, 800.0),
}


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


def _model(O, ph, fused):
    cfg = O.default_cfg(10, 1, ph=[ph[0], ph[1]])
    m = O.model_of(cfg)
    m.shock_fused = fused
    return m


def _problem(O, m, Y, vpre, p0, p1):
    L = O.lib()
    vH = [L.ora_hug_v(C.byref(m), k, p0, p1, vpre[k]) for k in range(2)]
    v0 = Y[0] * vpre[0] + Y[1] * vpre[1]
    vend = Y[0] * vH[0] + Y[1] * vH[1]
    return v0, vend, (p1 - p0) / (v0 - vend), vH


def _rhs_ld(ph, Y, p0, v0, j2, c, v, y):
    """the reference's right-hand side (PHY:302-310) in extended precision, and the scale of its terms"""
    g = [LD(ph[k][0]) for k in range(2)]
    pi = [LD(ph[k][1]) for k in range(2)]
    b = [LD(ph[k][5]) for k in range(2)]
    Yl = [LD(Y[0]), LD(Y[1])]
    p, v1, vv = LD(y[0]), LD(y[1]), LD(v)
    v2 = (vv - Yl[0] * v1) / Yl[1]
    P = LD(p0) + LD(j2) * (LD(v0) - vv)
    lam = [LD(c), LD(1.0) - LD(c)]
    vk = [v1, v2]
    num, den, sn = LD(-1.0), LD(0.0), LD(1.0)
    r, ep, evp = [0, 0], [0, 0], [0, 0]
    for k in range(2):
        ep[k] = (vk[k] - b[k]) / (g[k] - 1)
        evp[k] = g[k] * (p + pi[k]) / (g[k] - 1)
        r[k] = lam[k] / Yl[k] * (p - P)
        num += Yl[k] * r[k] / evp[k]
        sn += abs(Yl[k] * r[k] / evp[k])
        den += Yl[k] * ep[k] / evp[k]
    d0 = num / den
    d1 = (r[0] - ep[0] * d0) / evp[0]
    scale = (sn / abs(den), (abs(r[0]) + abs(ep[0] * d0)) / abs(evp[0]), abs(p) + abs(P))
    return (d0, d1, p - P), scale


@pytest.mark.parametrize("pair", list(PAIRS))
def test_fused_rhs_is_as_accurate_as_the_reference_arithmetic(O, pair):
    ph0, ph1, p0, p1, rho0, rho1 = PAIRS[pair]
    ph = (ph0, ph1)
    mr, mf = _model(O, ph, 0), _model(O, ph, 1)
    L = O.lib()
    rng = np.random.default_rng(11)
    eps = float(np.finfo(np.float64).eps)
    checked = 0
    worst = 0.0
    for _ in range(3000):
        Y = np.array([rng.uniform(0.05, 0.95), 0.0])
        Y[1] = 1.0 - Y[0]
        vpre = np.array([1.0 / (rho0 * rng.uniform(0.9, 1.1)), 1.0 / (rho1 * rng.uniform(0.9, 1.1))])
        pe = p0 + (p1 - p0) * rng.uniform(0.05, 1.0)
        v0, vend, j2, vH = _problem(O, mr, Y, vpre, p0, pe)
        if not (vend < v0):
            continue
        c = rng.uniform(0.01, 0.99)
        t = rng.uniform(0.0, 1.0)
        v = v0 + t * (vend - v0)
        y = np.array([p0 + (pe - p0) * rng.uniform(0.0, 1.2), vpre[0] + rng.uniform(0.0, 1.0) * (vH[0] - vpre[0]), 0.0])
        dr, df = np.zeros(3), np.zeros(3)
        okr = L.ora_prof_rhs(C.byref(mr), O._dp(Y), O._dp(vpre), p0, pe, c, v, O._dp(y), O._dp(dr))
        okf = L.ora_prof_rhs(C.byref(mf), O._dp(Y), O._dp(vpre), p0, pe, c, v, O._dp(y), O._dp(df))
        assert okr == okf, (pair, Y, vpre, pe, c, v, y)
        if not okr:
            continue
        exact, scale = _rhs_ld(ph, Y, p0, v0, j2, c, v, y)
        for k in range(3):
            er = float(abs(LD(dr[k]) - exact[k]))
            ef = float(abs(LD(df[k]) - exact[k]))
            tol = 4.0 * er + 16.0 * eps * float(scale[k])
            assert ef <= tol, (pair, k, ef, er, float(scale[k]), dr[k], df[k])
            worst = max(worst, ef / (eps * float(scale[k])))
        checked += 1
    assert checked > 1500, checked
    print(pair, checked, f"worst fused error {worst:.1f} ulp of the term scale")


def test_planted_term_swap_is_caught(O):
    """the comparison above fails for a fused right-hand side with e1 x1 in place of e1 x0 (planted in a Python copy of
    the fused formula, evaluated in double)"""
    ph0, ph1, p0, p1, rho0, rho1 = PAIRS["water_air"]
    mr = _model(O, (ph0, ph1), 0)
    rng = np.random.default_rng(3)
    eps = float(np.finfo(np.float64).eps)
    bad = 0
    for _ in range(200):
        Y = np.array([rng.uniform(0.05, 0.95), 0.0])
        Y[1] = 1.0 - Y[0]
        vpre = np.array([1.0 / rho0, 1.0 / rho1])
        v0, vend, j2, vH = _problem(O, mr, Y, vpre, p0, p1)
        c = rng.uniform(0.01, 0.99)
        v = v0 + rng.uniform(0, 1) * (vend - v0)
        y = np.array([p0 + (p1 - p0) * rng.uniform(0.0, 1.0), vpre[0] + rng.uniform(0.0, 1.0) * (vH[0] - vpre[0]), 0.0])
        v2 = (v - Y[0] * y[1]) / Y[1]
        if not (y[1] > 0 and v2 > 0):
            continue
        P = p0 + j2 * (v0 - v)
        dp = y[0] - P
        lr0, lr1 = c / Y[0], (1 - c) / Y[1]
        x0, x1 = ph0[0] * (y[0] + ph0[1]), ph1[0] * (y[0] + ph1[1])
        e0, e1 = Y[0] * (ph0[0] - 1) * lr0, Y[1] * (ph1[0] - 1) * lr1
        den = Y[0] * y[1] * x1 + Y[1] * v2 * x0
        num_wrong = -x0 * x1 + dp * (e0 * x1 + e1 * x1)
        exact, scale = _rhs_ld((ph0, ph1), Y, p0, v0, j2, c, v, y)
        if abs(num_wrong / den - float(exact[0])) > 16 * eps * float(scale[0]) + 1e-300:
            bad += 1
    assert bad > 150


def test_shares_agree_with_the_reference_arithmetic(O):
    L = O.lib()
    rng = np.random.default_rng(5)
    n = agree = 0
    for pair in PAIRS:
        ph0, ph1, p0, p1, rho0, rho1 = PAIRS[pair]
        mr, mf = _model(O, (ph0, ph1), 0), _model(O, (ph0, ph1), 1)
        for _ in range(40):
            Y = np.array([rng.uniform(0.02, 0.98), 0.0])
            Y[1] = 1.0 - Y[0]
            vpre = np.array([1.0 / (rho0 * rng.uniform(0.8, 1.2)), 1.0 / (rho1 * rng.uniform(0.8, 1.2))])
            pe = p0 * np.exp(rng.uniform(np.log(1.5), np.log(p1 / p0)))
            lr, lf, qr, qf = np.zeros(2), np.zeros(2), np.zeros(1), np.zeros(1)
            okr = L.ora_shock_share(C.byref(mr), O._dp(Y), O._dp(vpre), p0, pe, O._dp(lr), O._dp(qr))
            okf = L.ora_shock_share(C.byref(mf), O._dp(Y), O._dp(vpre), p0, pe, O._dp(lf), O._dp(qf))
            assert okr == okf, (pair, Y, vpre, pe)
            n += 1
            if okr:
                assert abs(lf[0] - lr[0]) <= 1e-9 and abs(lf[1] - lr[1]) <= 1e-9, (pair, lr, lf)
                assert abs(qf[0] - qr[0]) <= 1e-9 * qr[0], (pair, qr, qf)
                agree += 1
    assert agree > 100, (agree, n)


def _runs(O, cfg, U, tend, steps, cell):
    """the run in the reference arithmetic, in the fused form, and in the reference arithmetic from U with one cell's
    total energy one ulp higher (the run's own sensitivity to a rounding difference)"""
    out = []
    for fused, pert in ((0, 0), (1, 0), (0, 1)):
        cfg.shock_fused = fused
        V = U.reshape(-1, 8).copy()
        if pert:
            V[cell, 7] = np.nextafter(V[cell, 7], np.inf)
        o = O.Oracle(cfg)
        o.set_state(V.reshape(U.shape))
        assert o.run(tend, steps) == 0
        out.append((o.stats(), o.state().reshape(-1, 8).copy()))
    return out


def _check(runs):
    """the same counts, and per field a distance to the reference run at most 10x the one-ulp run's plus 1e-12 relative"""
    (s0, S0), (s1, S1), (_, S2) = runs
    for k in ("steps", "rk1", "rejects", "stages"):
        assert s0[k] == s1[k], (k, s0[k], s1[k])
    worst = []
    for f in range(8):
        ref = np.abs(S0[:, f]).sum()
        d_fused, d_ulp = np.abs(S1[:, f] - S0[:, f]).sum(), np.abs(S2[:, f] - S0[:, f]).sum()
        assert d_fused <= 10.0 * d_ulp + 1e-12 * ref, (f, d_fused / ref, d_ulp / ref)
        worst.append((d_fused / max(ref, 1e-300), d_ulp / max(ref, 1e-300)))
    print(" ".join(f"{a:.1e}/{b:.1e}" for a, b in worst))


def test_water_air_tube_runs_agree(O):
    cfg, U, tend = O.case_init(5, 200)
    _check(_runs(O, cfg, U, tend, -1, 100))


def test_shock_bubble_2d_runs_agree(O):
    from diffusionimex_amd import cases
    from tests.helpers import oracle_cfg_from
    scfg, U = cases.shock_bubble(40, 24, seed=3)
    _check(_runs(O, oracle_cfg_from(scfg, O), U, 1.0, 12, 500))
