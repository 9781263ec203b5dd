"""Extracts the reference's archived full-precision 1D profiles into tests/golden/ref_profiles.npz.

Source files (data, not code) under the reference tree: data/Schropff2026_Riemann_problems/c{80..88}_200.dat,
data/SPB2009_4.3.1_water_air/water_air_N100_t240us.dat (case 5), data/SPB2009_4.1_interface_advection/
interface_advection_t2.79ms.dat (case 4, N = 200). Columns kept (17 significant digits in the source): p u v rho alpha2
Y2 rho1 rho2 p1 p2 (write_profile, PC:1002-1013); and data/PS2014_water_cavitation_tube_{2mps,500mps}/ (cases 1 and 2,
5000 cells, relaxation levels 0 to 3), data/Pelanti2022_8.2.1_water_cavitation/ (case 30, levels 0 to 2),
data/PS2014_dodecane_shock_tube/ (case 3, 2000 cells, levels 0, 2, 3), data/Pelanti2022_8.2.2_dodecane_NASG/
(case 31, levels 0, 2), data/SPB2009_4.2_nitromethane/ (case 60) and data/Pelanti2022_8.1.1_JWL_water/ (case 61).
Run: python tests/golden/make_ref_profiles.py /root/reference
"""
import os
import sys

import numpy as np

COLS = {"p": 2, "u": 3, "v": 4, "rho": 5, "alpha2": 6, "Y2": 7, "rho1": 8, "rho2": 9, "p1": 14, "p2": 15}
FILES = {f"c{c}_200": (c, 200, f"data/Schropff2026_Riemann_problems/c{c}_200.dat") for c in range(80, 89)}
FILES["c5_100"] = (5, 100, "data/SPB2009_4.3.1_water_air/water_air_N100_t240us.dat")
FILES["c4_200"] = (4, 200, "data/SPB2009_4.1_interface_advection/interface_advection_t2.79ms.dat")
# Pelanti-Shyue water cavitation tubes (cases 1, 2; 5000 cells) at relaxation levels 0 (p), 1 (p-pT), 2 (p-pT-pTg) and
# 3 (p-pTg): the columns p u rho alpha2 Y2 p1 p2 (the others follow from them or carry fewer printed digits)
RELAX_COLS = ["p", "u", "rho", "alpha2", "Y2", "p1", "p2"]
RELAX_FILES = {"c1_5000_r0": (1, 5000, 0, "data/PS2014_water_cavitation_tube_2mps/p_relax.dat"),
               "c1_5000_r1": (1, 5000, 1, "data/PS2014_water_cavitation_tube_2mps/p-pT_relax.dat"),
               "c1_5000_r2": (1, 5000, 2, "data/PS2014_water_cavitation_tube_2mps/p-pT-pTg_relax.dat"),
               "c1_5000_r3": (1, 5000, 3, "data/PS2014_water_cavitation_tube_2mps/p-pTg_relax.dat"),
               "c2_5000_r1": (2, 5000, 1, "data/PS2014_water_cavitation_tube_500mps/p-pT_relax.dat"),
               "c2_5000_r2": (2, 5000, 2, "data/PS2014_water_cavitation_tube_500mps/p-pT-pTg_relax.dat"),
               # Pelanti's water cavitation tube (case 30, NASG liquid with covolume), levels 0-2
               "c30_5000_r0": (30, 5000, 0, "data/Pelanti2022_8.2.1_water_cavitation/p_relax.dat"),
               "c30_5000_r1": (30, 5000, 1, "data/Pelanti2022_8.2.1_water_cavitation/pT_relax.dat"),
               "c30_5000_r2": (30, 5000, 2, "data/Pelanti2022_8.2.1_water_cavitation/pTg_relax.dat"),
               # and its finite-rate runs: level 1 with theta = 2000, level 2 with nu = 1e-4 (meta: case, n, level, theta, nu)
               "c30_5000_theta2000": (30, 5000, 1, "data/Pelanti2022_8.2.1_water_cavitation/pT_finite_theta2000.dat"),
               "c30_5000_nu1e-4": (30, 5000, 2, "data/Pelanti2022_8.2.1_water_cavitation/pTg_finite_nu1e-4.dat"),
               # the dodecane liquid / vapour shock tubes (2000 cells): Pelanti-Shyue's SG form (case 3, material CFL
               # 0.05) at levels 0, 2 and 3, and Pelanti's NASG form (case 31) at levels 0 and 2
               "c3_2000_r0": (3, 2000, 0, "data/PS2014_dodecane_shock_tube/p_relax.dat"),
               "c3_2000_r2": (3, 2000, 2, "data/PS2014_dodecane_shock_tube/p-pT-pTg_relax.dat"),
               "c3_2000_r3": (3, 2000, 3, "data/PS2014_dodecane_shock_tube/p-pTg_relax.dat"),
               "c31_2000_r0": (31, 2000, 0, "data/Pelanti2022_8.2.2_dodecane_NASG/p_relax.dat"),
               "c31_2000_r2": (31, 2000, 2, "data/Pelanti2022_8.2.2_dodecane_NASG/pTg_relax.dat"),
               # the Mie-Grueneisen tubes (500 cells, level 0): Saurel-Petitpas-Berry's nitromethane (Cochran-Chan, case 60)
               # and Pelanti's detonation gases (JWL) | water (NASG) (case 61)
               "c60_500_r0": (60, 500, 0, "data/SPB2009_4.2_nitromethane/nitromethane_t67us.dat"),
               "c61_500_r0": (61, 500, 0, "data/Pelanti2022_8.1.1_JWL_water/JWL_water_t0.9ms.dat")}
RATES = {"c30_5000_theta2000": (2000.0, -1.0), "c30_5000_nu1e-4": (-1.0, 1e-4)}


def main(ref):
    out = {}
    for key, (case, n, rel) in FILES.items():
        a = np.loadtxt(os.path.join(ref, rel))
        assert a.shape[0] == n, (key, a.shape)
        for name, col in COLS.items():
            out[f"{key}/{name}"] = a[:, col]
        out[f"{key}/meta"] = np.array([case, n], dtype=np.int64)
    for key, (case, n, relax, rel) in RELAX_FILES.items():
        a = np.loadtxt(os.path.join(ref, rel))
        assert a.shape[0] == n, (key, a.shape)
        for name in RELAX_COLS:
            out[f"{key}/{name}"] = a[:, COLS[name]]
        out[f"{key}/meta"] = np.array([case, n, relax], dtype=np.int64)
        if key in RATES:
            out[f"{key}/rates"] = np.array(RATES[key])
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_profiles.npz")
    np.savez_compressed(dst, **out)
    print(dst, len(out))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
