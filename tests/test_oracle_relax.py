"""Relaxation levels 1-3 in the oracle: level 1 (p-pT: thermal equilibrium in the mixture cells, the pT closure
PHY:660-672 with its root finders PHY:183-225), levels 2 (p-pT-pTg) and 3 (p-pTg: chemical equilibrium where the phases
exceed the saturation temperature, the pTg closure PHY:756-778 on the saturation line T_sat(p) PHY:111-163, the
activation criteria PHY:814-830), and the step's RelaxationActivation PC:537-560, against the reference's archived
full-precision profiles of the Pelanti-Shyue water cavitation tubes (data/PS2014_water_cavitation_tube_2mps and _500mps,
5000 cells; extracted into tests/golden/ref_profiles.npz). In its PC mode the oracle reproduces them bit for bit (the
vapour mass fraction Y2, which the pTg closure changes, included), with the archived step, first-order-step and Newton
counts (data/*/note.md). The saturation line itself is checked against what thermodynamics fixes: equal Gibbs energies
at T_sat(p), the vapour branch (s2 > s1), and the Clausius-Clapeyron slope dp/dT = (s2 - s1) / (v2 - v1)."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_profiles.npz")
DIAG = {"p": 0, "u": 1, "rho": 3, "alpha2": 4, "Y2": 5, "p1": 12, "p2": 13}
NOTES = {"c1_5000_r0": (188, 33, 719), "c1_5000_r1": (188, 33, 695), "c1_5000_r2": (188, 33, 791),
         "c1_5000_r3": (188, 33, 883), "c2_5000_r1": (7252, 4, 14504), "c2_5000_r2": (7252, 4, 15990),
         "c30_5000_r0": (178, 33, 867), "c30_5000_r1": (178, 33, 666), "c30_5000_r2": (178, 33, 786),
         "c30_5000_theta2000": (178, 33, 866), "c30_5000_nu1e-4": (178, 33, 666),
         "c3_2000_r0": (2675, 3, 10719), "c3_2000_r2": (5896, 3, 23587), "c3_2000_r3": (5896, 3, 23587)}


def _check(O, gold, key, bound=0.0):
    case, n, relax = (int(v) for v in gold[key + "/meta"])
    cfg, U, tend = O.case_init(case, n)
    O.reference_mode(cfg, "PC")
    cfg.relax = relax
    if f"{key}/rates" in gold.files:
        cfg.theta_rate, cfg.nu_rate = (float(v) for v in gold[f"{key}/rates"])
    o = O.Oracle(cfg)
    o.set_state(U)
    assert o.run(tend) == 0
    o.L.ora_relax_activate(o.h)   # the reference's main activates once more before its output (PC:1300)
    D = O.cell_diag(cfg, o.state())
    for name, col in DIAG.items():
        a, b = D[:, col], gold[f"{key}/{name}"]
        if bound == 0.0:
            assert np.array_equal(a, b), f"{key}: column {name} differs in {np.count_nonzero(a != b)} cells"
        else:
            assert np.abs(a - b).sum() <= bound * np.abs(b).sum(), (key, name, np.abs(a - b).sum() / np.abs(b).sum())
    st = o.stats()
    assert (st["steps"], st["rk1"], st["newton_it"]) == NOTES[key]


@pytest.mark.parametrize("key", ["c1_5000_r0", "c1_5000_r1", "c1_5000_r2", "c1_5000_r3"])
def test_cavitation_tube_reproduced_bit_for_bit(oracle_lib, key):
    _check(oracle_lib, np.load(GOLD), key)


@pytest.mark.parametrize("key", ["c30_5000_r1", "c30_5000_r2", "c30_5000_theta2000", "c30_5000_nu1e-4"])
def test_nasg_cavitation_tube_reproduced_bit_for_bit(oracle_lib, key):
    """Pelanti's water cavitation tube (case 30: NASG liquid with covolume b = 6.61e-4, vapour): levels 1 and 2, and the
    finite rates (level 1 with theta = 2000: the theta closure, PHY:679-700; level 2 with nu = 1e-4: the nu closure,
    PHY:716-753; backward-Euler form), bit for bit with their Newton counts."""
    _check(oracle_lib, np.load(GOLD), key)


def test_nasg_cavitation_tube_level0(oracle_lib):
    """Case 30 at level 0 (every cell on the p closure with the shock-budget defect split): the steps, first-order steps
    and Newton count exactly, every column within 2e-9 relative L1 (measured 1.1e-11 in p, 4.7e-11 in u, 8.2e-10 in
    alpha2, in both the serial and the OpenMP program's paths). Parity unpinned below that: at levels 1-2 the same
    tube is bit-identical, so the rest lies in the level-0-only path (the p closure's defect split and the shock
    budget's shares with the covolume phase); its cause is not found."""
    _check(oracle_lib, np.load(GOLD), "c30_5000_r0", bound=2e-9)


@pytest.mark.slow
@pytest.mark.skipif(not os.environ.get("SPB6_SLOW"), reason="7252 steps of 5000 cells (about a minute); SPB6_SLOW=1 runs it")
@pytest.mark.parametrize("key", ["c2_5000_r1", "c2_5000_r2"])
def test_fast_cavitation_tube_reproduced_bit_for_bit(oracle_lib, key):
    _check(oracle_lib, np.load(GOLD), key)


def test_dodecane_shock_tube_level0_reproduced_bit_for_bit(oracle_lib):
    """Pelanti-Shyue's dodecane liquid / vapour shock tube (case 3: SG liquid at 1e8 Pa | vapour at 1e5 Pa, 2000 cells,
    material CFL 0.05, t = 473 us; 2675 steps): a strong shock into the vapour, so the strong-shock branch of the shock
    budget (the profile shooting of PHY:302-377) sets the phase shares of every shocked cell"""
    _check(oracle_lib, np.load(GOLD), "c3_2000_r0")


@pytest.mark.slow
@pytest.mark.skipif(not os.environ.get("SPB6_SLOW"), reason="5896 steps of 2000 cells (about a minute each); SPB6_SLOW=1 runs them")
@pytest.mark.parametrize("key", ["c3_2000_r2", "c3_2000_r3"])
def test_dodecane_shock_tube_levels23_reproduced_bit_for_bit(oracle_lib, key):
    _check(oracle_lib, np.load(GOLD), key)


def _variant_spread(O, gold, key):
    """relative L1 distance per column to the archive of the oracle in PC mode and of the reference's own variants: its
    OpenMP program's paths (libm), and its serial paths with the portable exp / log / pow; with the step counts"""
    case, n, relax = (int(v) for v in gold[key + "/meta"])
    out = {}
    for name, setup in (("PC", lambda c: O.reference_mode(c, "PC")), ("omp", lambda c: O.reference_mode(c, "PC_omp")),
                        ("port", lambda c: (O.reference_mode(c, "PC"), setattr(c, "portable_math", 1)))):
        cfg, U, tend = O.case_init(case, n)
        setup(cfg)
        cfg.relax = relax
        o = O.Oracle(cfg)
        o.set_state(U)
        assert o.run(tend) == 0
        o.L.ora_relax_activate(o.h)
        D = O.cell_diag(cfg, o.state())
        s = o.stats()
        out[name] = ({k: np.abs(D[:, c] - gold[f"{key}/{k}"]).sum() / np.abs(gold[f"{key}/{k}"]).sum() for k, c in DIAG.items()},
                     (s["steps"], s["rk1"], s["newton_it"]))
    return out


def test_dodecane_nasg_tube_level0(oracle_lib):
    """Pelanti's dodecane tube with the NASG liquid (case 31: covolume b = 7.51e-4; 2000 cells, 1144 steps) at level 0:
    the archived step, first-order-step and Newton counts exactly in all three paths, and every column within twice the
    larger distance of the reference's own variants to the archive, plus 1e-12. The run is sensitive to last-ulp changes:
    those variants land 2e-8 to 6e-7 from the archive (measured: p 3.2e-8 in PC mode, 6.6e-8 and 5.9e-8 in the variants;
    alpha2 up to 3.1e-7; Y2 up to 5.6e-6), so it is not bit for bit; a relative change of 1e-6 in the liquid covolume
    moves p by 1.5e-6 (25 times the bound), so the comparison pins the covolume terms. Parity of the last bits is
    unpinned, as for case 30 at level 0: the SG tube of case 3 is bit-identical through the same path."""
    O = oracle_lib
    sp = _variant_spread(O, np.load(GOLD), "c31_2000_r0")
    for name, (_, counts) in sp.items():
        assert counts == (1144, 5, 4575), (name, counts)
    for k in DIAG:
        assert sp["PC"][0][k] <= 2.0 * max(sp["omp"][0][k], sp["port"][0][k]) + 1e-12, (k, {n: v[0][k] for n, v in sp.items()})


@pytest.mark.slow
@pytest.mark.skipif(not os.environ.get("SPB6_SLOW"), reason="three runs of 1326 steps (about 35 s); SPB6_SLOW=1 runs it")
def test_dodecane_nasg_tube_level2(oracle_lib):
    """Case 31 at level 2 (pTg where the phases exceed saturation): the reference's variants already differ in their step
    counts (1326 to 1327 steps, Newton 5304 to 5309 against the archived 5305), so the run is held to the archive's step
    and first-order counts within one step and to twice the variants' spread per column; it pins little beyond that
    (a relative change of 1e-6 in the covolume stays inside the spread)."""
    O = oracle_lib
    sp = _variant_spread(O, np.load(GOLD), "c31_2000_r2")
    steps, rk1, _ = sp["PC"][1]
    assert abs(steps - 1326) <= 1 and rk1 == 5
    for k in DIAG:
        assert sp["PC"][0][k] <= 2.0 * max(sp["omp"][0][k], sp["port"][0][k]) + 1e-12, (k, {n: v[0][k] for n, v in sp.items()})


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


def _gibbs(ph, p, T):
    """g_k, s_k, v_k of a stiffened-gas phase (PHY:78-81) in extended precision"""
    g, pi, q, qp, cv, b = (np.longdouble(x) for x in ph)
    p, T = np.longdouble(p), np.longdouble(T)
    s = cv * (g * np.log(T) - (g - 1) * np.log(p + pi)) + qp
    return g * cv * T + b * p + q - T * s, s, (g - 1) * cv * T / (p + pi) + b


def test_saturation_line_thermodynamics(oracle_lib):
    """T_sat(p) of the water-vapour pair of the cavitation tubes: g1 = g2 at (p, T_sat) to round-off of g, the vapour
    branch (s2 > s1, v2 > v1), T_sat increasing, and the Clausius-Clapeyron slope dp/dT = (s2 - s1) / (v2 - v1) against
    central differences of T_sat (1e-5 relative); the table spans [0.1 Pa, pcrit) with pcrit where v1 = v2."""
    O = oracle_lib
    cfg, U, _ = O.case_init(1, 200)
    cfg.relax = 2
    o = O.Oracle(cfg)
    o.set_state(U)
    info = np.zeros(3)
    npts = o.L.ora_sat_info(o.h, O._dp(info))
    pmin, pcrit, vcrit = info
    assert npts > 1000 and abs(pmin - 0.1) < 1e-12 and 1e6 < pcrit < 5e9
    ph = [[cfg.ph[k][q] for q in range(6)] for k in range(2)]
    Tprev = 0.0
    for p in np.geomspace(1.0, 0.99 * pcrit, 60):
        T = o.L.ora_tsat(o.h, p)
        assert np.isfinite(T) and T > Tprev
        Tprev = T
        g1, s1, v1 = _gibbs(ph[0], p, T)
        g2, s2, v2 = _gibbs(ph[1], p, T)
        assert abs(float(g1 - g2)) <= 1e-11 * abs(float(g1)) + 1e-8, (p, T, float(g1 - g2))
        assert s2 > s1 and v2 > v1
        h = 1e-5 * p
        dTdp = (o.L.ora_tsat(o.h, p + h) - o.L.ora_tsat(o.h, p - h)) / (2 * h)
        assert abs(float((v2 - v1) / (s2 - s1)) / dTdp - 1) < 1e-5, (p, dTdp, float((v2 - v1) / (s2 - s1)))
    assert np.isnan(o.L.ora_tsat(o.h, 1.01 * pcrit))
    _, _, v1c = _gibbs(ph[0], pcrit, o.L.ora_tsat(o.h, pcrit * (1 - 1e-9)))
    assert abs(float(v1c) / vcrit - 1) < 1e-6


def test_level2_chemical_cells_sit_on_the_saturation_line(oracle_lib):
    """After 40 steps and an activation at level 2 the tube's rarefaction holds cells where vapour has formed (Y2 above
    twice its initial 5.5e-6): each is in chemical equilibrium, both phases at T_sat(p) and g1 = g2, to the round-off of
    recomputing p and T from the conserved state (the liquid's p_inf = 6e8 against p ~ 5e4 Pa costs four digits)."""
    O = oracle_lib
    cfg, U, tend = O.case_init(1, 400)
    cfg.relax = 2
    o = O.Oracle(cfg)
    o.set_state(U)
    Y0 = O.cell_diag(cfg, U)[:, 5].max()
    assert o.run(tend, 40) == 0
    o.L.ora_relax_activate(o.h)
    D = O.cell_diag(cfg, o.state())
    p, Y2, T1, T2 = D[:, 0], D[:, 5], D[:, 10], D[:, 11]
    ph = [[cfg.ph[k][q] for q in range(6)] for k in range(2)]
    chem = np.where(Y2 > 2 * Y0)[0]
    assert len(chem) > 5, (Y0, Y2.max())
    for i in chem:
        Ts = o.L.ora_tsat(o.h, p[i])
        assert abs(T1[i] / Ts - 1) < 1e-10 and abs(T2[i] / Ts - 1) < 1e-10 and abs(T1[i] / T2[i] - 1) < 1e-14, (i, T1[i], T2[i], Ts)
        g1, _, _ = _gibbs(ph[0], p[i], T1[i])
        g2, _, _ = _gibbs(ph[1], p[i], T1[i])
        assert abs(float(g1 - g2)) <= 1e-11 * abs(float(g1)), (i, float(g1 - g2), float(g1))


@pytest.mark.parametrize("relax", [2, 3])
def test_levels23_device_mode_spread_is_the_references_own(oracle_lib, relax):
    """What separates the device mode from the archive at levels 2-3. The phase-change decisions (T > T_sat(p), the
    Y clamps) flip on last-ulp changes: the serial archive is reproduced bit for bit by the oracle in the OpenMP
    program's iteration paths and the device's summation tree as long as libm's exp and log are used (level 2: exactly;
    level 3: the OpenMP paths alone already move it, by the reference's own PC / PC_omp spread); swapping in the
    portable exp and log of the device, in the serial program's paths, moves it by about as much as the device mode does.
    Measured (relative L1, p / alpha2 / Y2): level 2, portable in PC paths 1.2e-5 / 6.2e-3 / 1.3e-2, device mode
    1.2e-5 / 6.2e-3 / 1.3e-2; level 3, PC_omp 7.2e-6 / 7.0e-3 / 1.7e-2, device mode 7.5e-6 / 8.5e-3 / 2.5e-2."""
    O = oracle_lib
    gold = np.load(GOLD)
    key = f"c1_5000_r{relax}"

    def run(setup):
        cfg, U, tend = O.case_init(1, 5000)
        setup(cfg)
        cfg.relax = relax
        o = O.Oracle(cfg)
        o.set_state(U)
        assert o.run(tend) == 0
        o.L.ora_relax_activate(o.h)
        D = O.cell_diag(cfg, o.state())
        return {name: np.abs(D[:, col] - gold[f"{key}/{name}"]).sum() / np.abs(gold[f"{key}/{name}"]).sum()
                for name, col in DIAG.items()}

    libm_device_paths = run(lambda c: setattr(c, "portable_math", 0))
    portable_pc_paths = run(lambda c: (O.reference_mode(c, "PC"), setattr(c, "portable_math", 1)))
    omp = run(lambda c: O.reference_mode(c, "PC_omp"))
    device = run(lambda c: None)
    if relax == 2:
        assert max(libm_device_paths.values()) == 0.0 and max(omp.values()) == 0.0
    else:
        assert all(libm_device_paths[k] == omp[k] for k in DIAG)
    ref = {k: max(portable_pc_paths[k], omp[k]) for k in DIAG}
    for k in DIAG:
        assert device[k] <= 2.0 * ref[k] + 1e-12, (relax, k, device[k], ref)
