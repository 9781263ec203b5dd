"""Pins of the CPU oracle's functions by what the mathematics fixes: EOS consistency, the closure root and its energy
consistency, closed-form limits, the Hugoniot relation, the relaxed shock profile's end state, linear algebra against
dense solves, conservation, well-balancedness, and the portable transcendental functions against libm."""
import ctypes as C
import math

import numpy as np
import pytest


def _model(O, ph=((4.4, 6e8), (1.4, 0.0)), **kw):
    cfg = O.default_cfg(10, 1, ph=[[g, pi, 0, 0, 1000.0, 0] for g, pi in ph], **kw)
    return cfg, O.model_of(cfg)


def _eps(g, pi, a, m, p):
    return (p + g * pi) * a / (g - 1.0)


def test_p_mix_inverts_the_phase_energies(oracle_lib):
    """p_mix(alpha, m, sum_k eps_k(alpha_k, m_k, p)) = p for any common pressure p (A2.3)."""
    O = oracle_lib
    _, m = _model(O)
    L = O.lib()
    rng = np.random.default_rng(0)
    for _ in range(2000):
        a1 = rng.uniform(1e-8, 1 - 1e-8)
        m1, m2 = a1 * rng.uniform(500, 1500), (1 - a1) * rng.uniform(0.1, 100)
        p = 10 ** rng.uniform(4, 9.5)
        rhoe = _eps(4.4, 6e8, a1, m1, p) + _eps(1.4, 0.0, 1 - a1, m2, p)
        pm = L.ora_p_mix(C.byref(m), a1, m1, m2, rhoe)
        assert abs(pm - p) <= 1e-12 * (p + 6e8), (a1, p, pm)


def _closure_inputs(rng, ph):
    """consistent transported states: phases at a common pressure p0, strained by s, with an energy defect"""
    (g0, pi0), (g1, pi1) = ph
    a1 = rng.uniform(1e-7, 1 - 1e-7)
    m1, m2 = a1 * rng.uniform(200, 1500), (1 - a1) * rng.uniform(0.5, 60)
    p0 = 10 ** rng.uniform(4.5, 9.5)
    s = rng.uniform(-2e-3, 2e-3)
    atr = (a1 * (1 - s), (1 - a1) * (1 - s))
    etr = (_eps(g0, pi0, atr[0], m1, p0), _eps(g1, pi1, atr[1], m2, p0))
    rhoeC = etr[0] + etr[1] + rng.normal() * 1e-4 * (etr[0] + etr[1])
    Q = (abs(rng.normal()) * 1e-3 * abs(rhoeC), abs(rng.normal()) * 1e-3 * abs(rhoeC)) if rng.uniform() < 0.3 else (0.0, 0.0)
    pg = p0 * (1 + rng.uniform(-0.5, 0.5))
    return np.array([atr[0], atr[1], m1, m2, etr[0], etr[1], s, rhoeC, pg, Q[0], Q[1]])


@pytest.mark.parametrize("path", [0, 1])
def test_closure_root_bracketed_and_energy_consistent(oracle_lib, path):
    """The p closure returns the root of Phi(p) = sum alpha_k(p) - 1: Phi changes sign across it (brute-force bracket),
    it agrees with a 200-step bisection within the reference's stopping tolerance, and sum_k eps_k = (rho e)^C holds
    (energy consistency of the defect split, B5.2) because sum alpha^tr = 1 - s."""
    O = oracle_lib
    ph = ((4.4, 6e8), (1.4, 0.0))
    _, m = _model(O, ph)
    m.closure_path = path
    L = O.lib()
    rng = np.random.default_rng(7 + path)
    ok = 0
    for _ in range(400):
        x = _closure_inputs(rng, ph)
        out = np.zeros(8)
        if not L.ora_closure(C.byref(m), O._dp(x), O._dp(out)):
            continue
        ok += 1
        p = out[6]
        phi = lambda q: L.ora_closure_phi(C.byref(m), O._dp(x), q)
        plo = -6e8
        tol = 4e-16 * (abs(p) + abs(plo)) + 1e-300
        # bisection on a local bracket around the returned root (Phi need not be monotone over the whole admissible
        # range: the trace phase's alpha_k(p) can decrease or increase, and the defect split is piecewise affine)
        lo, hi = p - 1e-6 * (abs(p) + 6e8), p + 1e-6 * (abs(p) + 6e8)
        assert phi(lo) > 0 > phi(hi)
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if phi(mid) > 0:
                lo = mid
            else:
                hi = mid
        assert abs(p - 0.5 * (lo + hi)) <= 1e3 * tol + 1e-9 * abs(p), (x, p, lo, hi)
        assert phi(p - 64 * tol) >= 0 >= phi(p + 64 * tol)
        esum = out[4] + out[5]
        assert abs(esum - x[7]) <= 1e-11 * (abs(out[4]) + abs(out[5])), (esum, x[7])
        assert abs(out[0] + out[1] - 1.0) <= 1e-13
    assert ok > 300


def test_closure_ideal_gas_limit(oracle_lib):
    """Two identical ideal-gas phases at s = 0: p = (gamma - 1) (rho e)."""
    O = oracle_lib
    _, m = _model(O, ((1.4, 0.0), (1.4, 0.0)))
    L = O.lib()
    rng = np.random.default_rng(3)
    for _ in range(200):
        a1 = rng.uniform(0.01, 0.99)
        m1, m2 = rng.uniform(0.1, 2), rng.uniform(0.1, 2)
        rhoe = rng.uniform(1e3, 1e6)
        e1 = rhoe * a1
        x = np.array([a1, 1 - a1, m1, m2, e1, rhoe - e1, 0.0, rhoe, rhoe * 0.3, 0, 0])
        out = np.zeros(8)
        assert L.ora_closure(C.byref(m), O._dp(x), O._dp(out))
        assert abs(out[6] - 0.4 * rhoe) <= 4e-16 * 0.4 * rhoe * 8


def test_hugoniot_closed_form(oracle_lib):
    """hug_v satisfies each phase's Hugoniot e1 - e0 = -(p0 + p1)/2 (v1 - v0) with the SG energy e = (p + g pi) v/(g-1)."""
    O = oracle_lib
    _, m = _model(O)
    L = O.lib()
    for k, (g, pi) in enumerate(((4.4, 6e8), (1.4, 0.0))):
        for p0, p1, v0 in ((1e5, 1e9, 1e-3), (1e5, 3e5, 0.8), (2e8, 9e8, 1.1e-3)):
            v1 = L.ora_hug_v(C.byref(m), k, p0, p1, v0)
            e0, e1 = (p0 + g * pi) * v0 / (g - 1), (p1 + g * pi) * v1 / (g - 1)
            assert abs((e1 - e0) + 0.5 * (p0 + p1) * (v1 - v0)) <= 1e-12 * abs(e1)


def test_shock_share_weak_limit_and_profile_end_state(oracle_lib):
    """Weak shocks use the closed form of Theorem 8 (q = sum_k w_k dp^3 / 6, w_k = Y_k v_k G_k / K_k^2); just above the
    branch switch the integrated profile gives the same share and dissipation to O(dp), and the profile's end state lies
    on both phases' Hugoniots."""
    O = oracle_lib
    _, m = _model(O, ((1.4, 0.0), (1.67, 0.0)))
    L = O.lib()
    Y = np.array([0.6, 0.4])
    vpre = np.array([1 / 1.2, 1 / 0.17])
    p0 = 1e5
    K = np.array([1.4 * p0, 1.67 * p0])
    G = np.array([1.2, 1.335])
    lam = np.zeros(2)
    q = np.zeros(1)
    dp_w = 0.5e-2 * K.min()
    assert L.ora_shock_share(C.byref(m), O._dp(Y), O._dp(vpre), p0, p0 + dp_w, O._dp(lam), O._dp(q))
    w = Y * vpre * G / K ** 2
    assert abs(q[0] - w.sum() * dp_w ** 3 / 6) <= 1e-14 * q[0]
    assert abs(lam[0] - w[0] / w.sum()) <= 1e-15
    dp_s = 2e-2 * K.min()
    lam2, q2 = np.zeros(2), np.zeros(1)
    assert L.ora_shock_share(C.byref(m), O._dp(Y), O._dp(vpre), p0, p0 + dp_s, O._dp(lam2), O._dp(q2))
    q_weak = w.sum() * dp_s ** 3 / 6
    assert abs(q2[0] / q_weak - 1) < 0.1
    assert abs(lam2[0] - w[0] / w.sum()) < 0.05
    # end state of the profile with the converged share: phase 1 ends on its Hugoniot (the shooting target) and, since
    # the mixture volume is fixed, phase 2 too
    v1e, qe = np.zeros(1), np.zeros(1)
    assert L.ora_prof_integrate(C.byref(m), O._dp(Y), O._dp(vpre), p0, p0 + dp_s, lam2[0], O._dp(v1e), O._dp(qe))
    vH = [L.ora_hug_v(C.byref(m), k, p0, p0 + dp_s, vpre[k]) for k in range(2)]
    assert abs(v1e[0] - vH[0]) / vpre[0] < 1e-9
    vend = Y[0] * vH[0] + Y[1] * vH[1]
    v2e = (vend - Y[0] * v1e[0]) / Y[1]
    assert abs(v2e - vH[1]) / vpre[1] < 1e-8
    assert abs(qe[0] - q2[0]) <= 1e-12 * q2[0]


def _ctx_2d(O, nx, ny, bc=(0, 0, 0, 0)):
    cfg = O.default_cfg(nx, ny, bc=list(bc), ph=[[1.4, 0, 0, 0, 1000.0, 0], [1.4, 0, 0, 0, 1000.0, 0]])
    return O.Oracle(cfg)


@pytest.mark.parametrize("order", [0, 1, 2])
def test_bicgstab_and_thomas_against_dense(oracle_lib, order):
    """BiCGStab (Jacobi + coarse-space preconditioner) and Thomas solve small diagonally dominant systems with the
    5-point / 3-point patterns to the tolerance the reference uses (1e-13 |b|), compared with numpy's dense solve."""
    O = oracle_lib
    rng = np.random.default_rng(11 + order)
    import ctypes
    for nx, ny, thomas in ((12, 9, 0), (40, 1, 0), (40, 1, 1)):
        cfg = O.default_cfg(nx, ny, dot_order=order, ph=[[1.4, 0, 0, 0, 1000.0, 0], [1.4, 0, 0, 0, 1000.0, 0]])
        o = O.Oracle(cfg)
        n = nx * ny
        val = np.zeros((n, 5))
        val[:, 1:3] = rng.uniform(-1, 0, size=(n, 2))
        if ny > 1:
            val[:, 3:5] = rng.uniform(-1, 0, size=(n, 2))
        val[:, 0] = np.abs(val[:, 1:]).sum(1) + rng.uniform(0.5, 2, size=n)
        A = np.zeros((n, n))
        for r in range(n):
            i, j = r % nx, r // nx
            cols = [r, r - 1 if i > 0 else r, r + 1 if i < nx - 1 else r]
            if ny > 1:
                cols += [r - nx if j > 0 else r, r + nx if j < ny - 1 else r]
            for q, c in enumerate(cols):
                A[r, c] += val[r, q]
        b = rng.normal(size=n)
        x = np.zeros(n)
        it = np.zeros(1, dtype=np.int32)
        rc = o.L.ora_linsolve(o.h, O._dp(val.ravel().copy()), O._dp(b), O._dp(x), thomas, it.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        assert rc == 0
        xd = np.linalg.solve(A, b)
        assert np.abs(x - xd).max() <= 1e-11 * np.abs(xd).max()


def test_fixed_order_sums(oracle_lib):
    """The three summation orders agree with an exactly rounded sum (math.fsum) to within n ulps; the device order is
    the documented tree (1024-term blocks, 4 strided terms per thread, xor butterflies, (w0 + w1) + (w2 + w3))."""
    O = oracle_lib
    L = O.lib()
    rng = np.random.default_rng(5)
    for n in (1, 7, 1023, 1024, 1025, 5000, 300000):
        a = rng.normal(size=n) * 10 ** rng.uniform(-3, 3, size=n)
        exact = math.fsum(a)
        for order in (0, 1, 2):
            s = L.ora_sum(order, O._dp(a), n)
            assert abs(s - exact) <= 2e-16 * n * np.abs(a).sum() + 1e-300
    # explicit tree order on a small case
    a = rng.normal(size=2048)

    def tree(v):
        w = []
        for q in range(4):
            x = list(v[q * 64:(q + 1) * 64])
            off = 32
            while off >= 1:
                x = [x[l] + x[l ^ off] for l in range(64)]
                off //= 2
            w.append(x[0])
        return (w[0] + w[1]) + (w[2] + w[3])
    parts = []
    for b in range(2):
        v = []
        for t in range(256):
            s = 0.0
            for it in range(4):
                s += a[b * 1024 + it * 256 + t]
            v.append(s)
        parts.append(tree(v))
    v = [0.0] * 256
    v[0] = 0.0 + parts[0]
    v[1] = 0.0 + parts[1]
    assert L.ora_sum(2, O._dp(a), 2048) == tree(v)


def test_portable_math_against_libm(oracle_lib):
    O = oracle_lib
    L = O.lib()
    rng = np.random.default_rng(9)

    def ulp(x):
        return math.ulp(x) if x != 0 else 5e-324
    for x in np.concatenate([rng.uniform(-40, 40, 3000), rng.uniform(-700, 700, 1000)]):
        assert abs(L.ora_pexp(x) - math.exp(x)) <= 3 * ulp(math.exp(x))
    for x in np.concatenate([10 ** rng.uniform(-300, 300, 3000), rng.uniform(0.5, 2.0, 1000)]):
        assert abs(L.ora_plog(x) - math.log(x)) <= 3 * ulp(math.log(x)) + 1e-300
    for x in 10 ** rng.uniform(-30, 30, 2000):
        for y in (0.25, 0.2):
            # exp(y log x): log's last-ulp error is amplified by |y log x|
            assert abs(L.ora_ppow(x, y) - x ** y) <= (4 + 2 * abs(y * math.log(x))) * ulp(x ** y)


@pytest.mark.parametrize("bc", [1, 2])
def test_conservation_closed_box(oracle_lib, bc):
    """2D box with a pressure bump, walls (bc 1) or periodic (bc 2) on every side: sum m1, sum m2 and sum E conserved to
    round-off (PC:1299); in the periodic box also both momenta (walls exert pressure forces, so there they change)."""
    O = oracle_lib
    nx, ny = 24, 16
    ph = [[4.4, 6e8, 0, 0, 1000.0, 0], [1.4, 0, 0, 0, 1000.0, 0]]
    cfg = O.default_cfg(nx, ny, bc=[bc] * 4, ph=ph, cfl=0.5)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)
    a1 = np.where((X - 0.4) ** 2 + (Y - 0.5) ** 2 < 0.05, 1e-6, 1 - 1e-6)
    p = 1e5 + 4e5 * np.exp(-((X - 0.6) ** 2 + (Y - 0.4) ** 2) / 0.01)
    m1, m2 = a1 * 1000.0, (1 - a1) * 1.0
    U = np.zeros((ny, nx, 8))
    U[..., 0], U[..., 1], U[..., 2] = a1, m1, m2
    U[..., 5] = (p + 4.4 * 6e8) * a1 / 3.4
    U[..., 6] = p * (1 - a1) / 0.4
    U[..., 7] = U[..., 5] + U[..., 6]
    o = O.Oracle(cfg)
    o.set_state(U)
    assert o.run(1.0, 8) == 0
    U1 = o.state()
    for k in (1, 2, 7):
        assert abs(U1[..., k].sum() - U[..., k].sum()) <= 2e-15 * nx * ny * np.abs(U[..., k]).max()
    if bc == 2:
        for k in (3, 4):   # the momenta start at zero: their sums stay zero to round-off of the moved momentum
            assert np.abs(U1[..., k]).max() > 0
            assert abs(U1[..., k].sum()) <= 2e-15 * nx * ny * np.abs(U1[..., k]).max()


def test_well_balanced_interface_advection(oracle_lib):
    """Case 4 (SPB 4.1): a water-air contact advected at 100 m/s keeps p and u uniform with zero Newton iterations."""
    O = oracle_lib
    cfg, U, tend = O.case_init(4, 200)
    o = O.Oracle(cfg)
    o.set_state(U)
    assert o.run(tend) == 0
    D = O.cell_diag(cfg, o.state())
    assert o.stats()["newton_it"] == 0
    assert np.abs(D[:, 0] / 1e5 - 1).max() < 1.5e-9
    assert np.abs(D[:, 1] / 100 - 1).max() < 1.5e-9
