"""Extracts the reference's archived 2D snapshot of case 70 (SPB 2009 4.6, Mach 1.5 shock in air onto a helium bubble,
600 x 160 cells on 0.3 x 0.08 m, PC run to t = 134 us) into tests/golden/case70_t134us.npz.

Source (data, not code): data/SPB2009_4.6_shock_helium_bubble/shock_helium_t134us.dat.gz under the reference tree,
columns x y p u v rho alpha2 Y2 T1 T2 at 9 significant digits (%.8e), one line per cell with x fastest. Kept: p u v rho alpha2
Y2 of every second cell in x and in y (cells (2i, 2j), 300 x 80 values per column), as float64.
Run: python tests/golden/make_case70.py /root/reference
"""
import gzip
import os
import sys

import numpy as np

COLS = ("p", "u", "v", "rho", "alpha2", "Y2")


def main(ref):
    path = os.path.join(ref, "data/SPB2009_4.6_shock_helium_bubble/shock_helium_t134us.dat.gz")
    a = np.loadtxt(gzip.open(path)).reshape(160, 600, 10)
    sub = a[::2, ::2]
    out = {name: np.ascontiguousarray(sub[:, :, 2 + k]) for k, name in enumerate(COLS)}
    out["meta"] = np.array([70, 600, 160, 2], dtype=np.int64)   # case, nx, ny, stride
    out["t"] = np.array(134e-6)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "case70_t134us.npz")
    np.savez_compressed(dst, **out)
    print(dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
