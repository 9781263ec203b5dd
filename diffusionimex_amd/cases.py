"""Initial states of the workloads (the caller side of the stepper: main's u_initialization, PC:450-465).

Each builder returns (Config, U[ny, nx, 8]) with U in the reference's conserved variables
(alpha1, m1, m2, rho u, rho v, eps1, eps2, E) built by prim2cell (PHY:849-854) in the same operation order.
Random draws use SplitMix64 seeded per workload.
"""
import math

import numpy as np

from .spb6 import make_config, BC_TRANSMISSIVE, BC_WALL

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014): 64-bit state, golden-gamma increment, two xor-shift-multiply rounds."""

    def __init__(self, seed):
        self.s = seed & MASK64

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, a=0.0, b=1.0):
        return a + (b - a) * ((self.next() >> 11) * (1.0 / (1 << 53)))


def eps_amp(g, pi, a, m, p, b=0.0, q=0.0):
    """phase energy eps_k(alpha, m, p) (PHY:69-71)."""
    return (p + g * pi) * (a - m * b) / (g - 1.0) + m * q


def prim2cell(ph, a1, r1, r2, u, v, p1, p2):
    """conserved state from primitives (PHY:849-854); arrays broadcast. ph[k] = (gamma, p_inf[, q, q', cv, b])."""
    a1 = np.asarray(a1, dtype=np.float64)
    a2 = 1 - a1
    m1 = a1 * r1
    m2 = a2 * r2
    rho = m1 + m2
    (g0, pi0), (g1, pi1) = ph[0][:2], ph[1][:2]
    q0, q1 = (ph[0][2] if len(ph[0]) > 2 else 0.0), (ph[1][2] if len(ph[1]) > 2 else 0.0)
    b0, b1 = (ph[0][5] if len(ph[0]) > 5 else 0.0), (ph[1][5] if len(ph[1]) > 5 else 0.0)
    U = np.empty(np.broadcast(a1, r1, r2, u, v, p1, p2).shape + (8,))
    U[..., 0] = a1
    U[..., 1] = m1
    U[..., 2] = m2
    U[..., 3] = rho * u
    U[..., 4] = rho * v
    U[..., 5] = eps_amp(g0, pi0, a1, m1, p1, b0, q0)
    U[..., 6] = eps_amp(g1, pi1, a2, m2, p2, b1, q1)
    U[..., 7] = U[..., 5] + U[..., 6] + 0.5 * rho * (u * u + v * v)
    return U


def water_air_tube(nx=1000, cfl=0.5, newton_tol=1e-10, **kw):
    """1D water-air shock tube: [0, 1] m, interface 0.7; left alpha_water = 1 - 1e-8, p = 1e9 Pa; right alpha_water = 1e-8,
    p = 1e5 Pa; rho_water = 1000, rho_air = 50 kg/m3; u = 0; SG water (4.4, 6e8), air (1.4, 0); transmissive."""
    ph = ((4.4, 6e8), (1.4, 0.0))
    cfg = make_config(nx, 1, phases=ph, cfl=cfl, newton_tol=newton_tol, **kw)
    x = (np.arange(nx) + 0.5) / nx
    L = x < 0.7
    U = prim2cell(ph, np.where(L, 1 - 1e-8, 1e-8), 1000.0, 50.0, 0.0, 0.0, np.where(L, 1e9, 1e5), np.where(L, 1e9, 1e5))
    return cfg, U.reshape(1, nx, 8)


def shock_bubble(nx=2048, ny=1024, seed=0, cfl=0.5, **kw):
    """2D shock-bubble on 0.2 x 0.1 m: ambient gas (gamma 1.4, rho 1.2, p 1e5); planar Mach-1.22 shock at x = 0.1 m moving
    +x with the Rankine-Hugoniot post-shock state on its left; light-gas bubble (gamma 1.648, rho 0.166, r = 0.025 m) with
    its centre drawn from the seed in x in [0.125, 0.175], y in [0.025, 0.075]; tracer alpha 1e-6 of the absent phase;
    reflective top/bottom walls, transmissive left/right. Cell-centre sampling (no sub-cell averaging)."""
    ph = ((1.4, 0.0), (1.648, 0.0))
    cfg = make_config(nx, ny, phases=ph, domain=(0.0, 0.2, 0.0, 0.1), bc=(BC_TRANSMISSIVE, BC_TRANSMISSIVE, BC_WALL, BC_WALL),
                      cfl=cfl, **kw)
    rng = SplitMix64(seed)
    xc, yc = rng.uniform(0.125, 0.175), rng.uniform(0.025, 0.075)
    g, Ms, p0, r0 = 1.4, 1.22, 1e5, 1.2
    c0 = math.sqrt(g * p0 / r0)
    p1 = p0 * (1 + 2 * g / (g + 1) * (Ms * Ms - 1))
    r1 = r0 * (g + 1) * Ms * Ms / ((g - 1) * Ms * Ms + 2)
    u1 = Ms * c0 * (1 - r0 / r1)      # post-shock gas velocity of a shock moving +x into gas at rest
    dx, dy = 0.2 / nx, 0.1 / ny
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dy
    X, Y = np.meshgrid(x, y)
    inb = (X - xc) ** 2 + (Y - yc) ** 2 < 0.025 ** 2
    post = X < 0.1
    a1 = np.where(inb, 1e-6, 1 - 1e-6)
    rg = np.where(post, r1, r0)
    u = np.where(post, u1, 0.0)
    p = np.where(post, p1, p0)
    U = prim2cell(ph, a1, rg, 0.166, u, 0.0, p, p)
    return cfg, U


def cavitation_2d(nx=4096, ny=4096, seed=0, cfl=0.5, **kw):
    """2D low-Mach cavitation on [0, 1]^2: water (4.4, 6e8), rho 1000, with gas volume fraction 1e-2 (gamma 1.4, rho 1);
    p0 = 1e5 Pa; u = 100 sign(x - 0.5), v = 100 sign(y - 0.5), each times (1 + 0.01 xi) with xi ~ U(-1, 1) per cell from
    the seed; transmissive boundaries."""
    ph = ((4.4, 6e8), (1.4, 0.0))
    cfg = make_config(nx, ny, phases=ph, cfl=cfl, **kw)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)
    rs = np.random.Generator(np.random.PCG64(SplitMix64(seed).next()))
    xi1 = rs.uniform(-1.0, 1.0, size=X.shape)
    xi2 = rs.uniform(-1.0, 1.0, size=X.shape)
    u = 100.0 * np.sign(X - 0.5) * (1 + 0.01 * xi1)
    v = 100.0 * np.sign(Y - 0.5) * (1 + 0.01 * xi2)
    U = prim2cell(ph, 1 - 1e-2, 1000.0, 1.0, u, v, 1e5, 1e5)
    return cfg, U


def perturbed_interface(nx=4096, ny=2048, seed=0, cfl=0.5, **kw):
    """2D Mach-1.5 (in gas) shock on a perturbed water-gas interface, [0, 2] x [0, 1]: interface x0 + A sin(2 pi k y / Ly)
    with x0 = 1.0, A ~ U[0.01, 0.05] Ly and k ~ U{1..8} from the seed; water (4.4, 6e8) rho 1000 right of it, gas
    (1.4, 0) rho 1 on the left with the shock at x = 0.5 moving +x; tracer 1e-6; walls top/bottom."""
    ph = ((4.4, 6e8), (1.4, 0.0))
    cfg = make_config(nx, ny, phases=ph, domain=(0.0, 2.0, 0.0, 1.0), bc=(BC_TRANSMISSIVE, BC_TRANSMISSIVE, BC_WALL, BC_WALL),
                      cfl=cfl, **kw)
    rng = SplitMix64(seed)
    A = rng.uniform(0.01, 0.05)
    k = 1 + int(rng.uniform(0.0, 8.0))
    dx, dy = 2.0 / nx, 1.0 / ny
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dy
    X, Y = np.meshgrid(x, y)
    g, Ms, p0, r0 = 1.4, 1.5, 1e5, 1.0
    c0 = math.sqrt(g * p0 / r0)
    p1 = p0 * (1 + 2 * g / (g + 1) * (Ms * Ms - 1))
    r1 = r0 * (g + 1) * Ms * Ms / ((g - 1) * Ms * Ms + 2)
    u1 = Ms * c0 * (1 - r0 / r1)
    water = X > 1.0 + A * np.sin(2 * math.pi * k * Y)
    post = X < 0.5
    a1 = np.where(water, 1 - 1e-6, 1e-6)
    U = prim2cell(ph, a1, 1000.0, np.where(post, r1, r0), np.where(post, u1, 0.0), 0.0,
                  np.where(post, p1, p0), np.where(post, p1, p0))
    return cfg, U


def riemann_batch(n=4096, seed=0, lmin=256, lmax=8192, cfl=0.5, newton_tol=1e-10, **kw):
    """Batched 1D two-phase Riemann problems on [0, 1] m, transmissive: lengths uniform in [lmin, lmax] cells; interface
    position uniform in [0.3, 0.7]; liquid left (alpha_liquid = 1 - 1e-6, rho 1000, p_L) and gas right (alpha_liquid =
    1e-6, rho 50, p_R), u = 0; p_L, p_R log-uniform in [1e5, 1e9] Pa; liquid stiffened gas gamma ~ U[2, 6],
    p_inf ~ U[3e8, 8e8] Pa; gas gamma 1.4. Returns (scheme Config, [BatchProblem], [U_p[nx_p, 8]], draws)."""
    from .batch import problem
    rng = SplitMix64(seed)
    probs, Us, draws = [], [], []
    for _ in range(n):
        nx = lmin + int(rng.uniform(0.0, lmax - lmin + 1))
        nx = min(nx, lmax)
        xi = rng.uniform(0.3, 0.7)
        pL = 10.0 ** rng.uniform(5.0, 9.0)
        pR = 10.0 ** rng.uniform(5.0, 9.0)
        gl = rng.uniform(2.0, 6.0)
        pil = rng.uniform(3e8, 8e8)
        ph = ((gl, pil), (1.4, 0.0))
        x = (np.arange(nx) + 0.5) / nx
        L = x < xi
        U = prim2cell(ph, np.where(L, 1 - 1e-6, 1e-6), 1000.0, 50.0, 0.0, 0.0, np.where(L, pL, pR), np.where(L, pL, pR))
        probs.append(problem(nx, ph))
        Us.append(U.reshape(nx, 8))
        draws.append({"nx": nx, "x_interface": xi, "pL": pL, "pR": pR, "gamma_liquid": gl, "pinf_liquid": pil})
    scheme = make_config(2, 1, cfl=cfl, newton_tol=newton_tol, **kw)
    return scheme, probs, Us, draws
