"""spb6_linsolve_device, the allocation-free form of spb6_linsolve (its work vectors in the library's fixed device array),
against spb6_linsolve on the same systems, bit for bit: two BiCGStab calls in a row, the second smaller (the array still
holds the first call's values), a Thomas call, and a size above the array's maximum (rejected before any work). The
test, as the caller, allocates the device buffers of its inputs and outputs with the HIP runtime the library uses."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(nx, ny, seed):
    """a diagonally dominant 5-point (2D) / 3-point (1D) system in spb6_linsolve's row layout (self, W, E, S, N)"""
    rng = np.random.default_rng(seed)
    n = nx * ny
    val = np.zeros((n, 5))
    val[:, 1:5 if ny > 1 else 3] = -rng.uniform(0.1, 1.0, size=(n, 4 if ny > 1 else 2))
    val[:, 0] = np.abs(val[:, 1:]).sum(axis=1) * rng.uniform(1.05, 1.6, size=n)
    b = rng.normal(size=n)
    return val, b


class _Hip:
    """the caller's side: device buffers from the HIP runtime the library is linked with"""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), a.nbytes) == 0
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0
        self.bufs.append(p)
        return p

    def get(self, p, n):
        out = np.empty(n)
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _device_solve(L, nx, ny, val, b, thomas=False):
    n = nx * ny
    H = _Hip()
    try:
        Jd, bd = H.put(val.T), H.put(b)      # J[w * n + r]
        xd = H.put(np.full(n, np.nan))
        bc = (C.c_int * 4)(0, 0, 0, 0)
        it = C.c_int(-1)
        rc = L.spb6_linsolve_device(nx, ny, bc, Jd, bd, xd, 1e-13, 4000, int(thomas), C.byref(it), None)
        return rc, H.get(xd, n), it.value
    finally:
        H.free()


def test_device_linsolve_matches_linsolve_twice_then_smaller():
    from diffusionimex_amd import spb6
    L = spb6.load()
    for nx, ny, seed in ((300, 200, 1), (120, 70, 2)):
        val, b = _system(nx, ny, seed)
        cfg = spb6.make_config(nx, ny)
        st = spb6.Stepper(cfg)
        x0, it0, rc0 = st.linsolve(val, b)
        st.close()
        rc, x, it = _device_solve(L, nx, ny, val, b)
        assert rc0 == 0 and rc == 0, (rc0, rc, spb6.load().spb6_last_error())
        assert it == it0 and it > 2
        assert np.array_equal(x.view(np.uint64), x0.view(np.uint64)), f"{nx}x{ny}: {np.count_nonzero(x != x0)} entries differ"


def test_device_thomas_matches_linsolve_and_rejects_oversize():
    from diffusionimex_amd import spb6
    L = spb6.load()
    val, b = _system(500, 1, 3)
    st = spb6.Stepper(spb6.make_config(500, 1))
    x0, _, rc0 = st.linsolve(val, b, thomas=True)
    st.close()
    rc, x, _ = _device_solve(L, 500, 1, val, b, thomas=True)
    assert rc0 == 0 and rc == 0
    assert np.array_equal(x.view(np.uint64), x0.view(np.uint64))
    bc = (C.c_int * 4)(0, 0, 0, 0)
    it = C.c_int(0)
    rc = L.spb6_linsolve_device(1024, 1024, bc, C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), 1e-13, 10, 0, C.byref(it), None)
    assert rc == spb6.STATUS_CODE["BAD_ARG"] and b"maximum" in L.spb6_last_error()
