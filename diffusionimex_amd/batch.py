"""Batched 1D problems on the GPU: many independent runs of the reference's 1D SPB6_PC_IMEX path (PC:1230-1296), each
with its own length, domain, boundaries and phases, advanced together (spb6_batch_* in include/spb6_imex.h)."""
import ctypes as C

import numpy as np

from . import spb6
from .spb6 import Config, Phase, Stats, SPB6Error, STATUS


class BatchProblem(C.Structure):
    _fields_ = [("nx", C.c_int), ("x0", C.c_double), ("x1", C.c_double), ("bc", C.c_int * 2), ("bcu", C.c_double * 2),
                ("ph", Phase * 2)]


class Tune(C.Structure):
    _fields_ = [("threads", C.c_int), ("per_cu", C.c_int), ("best_ms", C.c_double), ("ntried", C.c_int),
                ("tried_threads", C.c_int * 16), ("tried_per_cu", C.c_int * 16), ("tried_ms", C.c_double * 16)]


def problem(nx, phases, domain=(0.0, 1.0), bc=(0, 0), bcu=(0.0, 0.0)):
    """One problem: nx cells on domain, boundaries (left, right), phases ((gamma, p_inf[, q, qp, cv, b]), ...)."""
    p = BatchProblem()
    p.nx = int(nx)
    p.x0, p.x1 = domain
    for s in range(2):
        p.bc[s] = bc[s]
        p.bcu[s] = bcu[s]
    ref = spb6.make_config(2, 1, phases=phases)   # phase defaults (cv) as in make_config
    for k in range(2):
        p.ph[k] = ref.ph[k]
    return p


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Batch:
    """A device-resident batch; `scheme` is a spb6 Config whose rk, CFL, ramp, tolerances, limiter and order apply to
    every problem (its grid, boundary and phase fields are ignored)."""

    def __init__(self, problems, scheme, device=0):
        self.L = spb6.load()
        self.np = len(problems)
        arr = (BatchProblem * self.np)(*problems)
        self.nx = np.array([p.nx for p in problems], dtype=np.int64)
        self.off = np.concatenate([[0], np.cumsum(self.nx)])
        self.h = self.L.spb6_batch_create(C.cast(arr, C.c_void_p), self.np, C.byref(scheme), int(device))
        if not self.h:
            raise SPB6Error("spb6_batch_create failed (bad problem or scheme, or out of device memory)")

    def close(self):
        if getattr(self, "h", None):
            self.L.spb6_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise SPB6Error(f"{what}: {STATUS.get(rc, rc)} {self.L.spb6_batch_last_error(self.h).decode()}")

    @property
    def cells(self):
        return int(self.off[-1])

    def set_state(self, U):
        """U: list of per-problem [nx, 8] arrays, or one [sum nx, 8] array."""
        if isinstance(U, (list, tuple)):
            U = np.concatenate([np.asarray(u, dtype=np.float64).reshape(-1, 8) for u in U])
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(-1, 8)
        assert U.shape[0] == self.cells
        self._check(self.L.spb6_batch_set_state(self.h, _dp(U)), "batch_set_state")

    def state(self):
        U = np.empty((self.cells, 8))
        self._check(self.L.spb6_batch_get_state(self.h, _dp(U)), "batch_get_state")
        return U

    def problem_state(self, U, p):
        return U[self.off[p]:self.off[p + 1]]

    def pressure(self):
        P = np.empty(self.cells)
        pref = np.empty(self.np)
        self._check(self.L.spb6_batch_get_pressure(self.h, _dp(P), _dp(pref)), "batch_get_pressure")
        return P, pref

    def run(self, t_end=1.0, steps=1):
        return self.L.spb6_batch_run(self.h, float(t_end), int(steps))

    def stats(self):
        per = (Stats * self.np)()
        tot = Stats()
        self._check(self.L.spb6_batch_get_stats(self.h, per, C.byref(tot)), "batch_get_stats")
        conv = lambda s: {k: getattr(s, k) for k, _ in Stats._fields_}
        return [conv(s) for s in per], conv(tot)

    def set_launch(self, threads, per_cu):
        self._check(self.L.spb6_batch_set_launch(self.h, int(threads), int(per_cu)), "batch_set_launch")

    def autotune(self, steps=4):
        t = Tune()
        self._check(self.L.spb6_batch_autotune(self.h, int(steps), C.cast(C.byref(t), C.c_void_p)), "batch_autotune")
        tried = [(t.tried_threads[k], t.tried_per_cu[k], t.tried_ms[k]) for k in range(t.ntried)]
        return {"threads": t.threads, "per_cu": t.per_cu, "best_ms": t.best_ms, "tried": tried}
