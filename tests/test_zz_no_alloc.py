"""tdt_monitor_noalloc (no context, no device allocation) against tdt_monitor.

Both compute the per-iteration monitor of turbnstwod.f90:147-168.  With the
context in grid.in order (TDT_ORDER_INPUT) the two sum the same cells in the
same blocks, so the norms must agree bit for bit.  Two calls in a row, the
second on a smaller grid: the static partial records the first call left must
not leak into the second.  Device buffers come from the HIP runtime libtdt
itself loaded (the caller's memory, as for a real caller of this entry point).
"""
import ctypes

import numpy as np
import pytest

import tdt
from tdt import lib as tlib

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    """The libamdhip64 that libtdt loaded (same runtime, same device memory)."""
    tlib.load()
    path = None
    with open("/proc/self/maps") as fh:
        for ln in fh:
            if "libamdhip64.so" in ln:
                path = ln.split()[-1]
                break
    assert path, "libamdhip64 not loaded by libtdt"
    H = ctypes.CDLL(path)
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    H.hipFree.argtypes = [ctypes.c_void_p]
    return H


class DevBuf:
    def __init__(self, H, a):
        self.H, self.nbytes = H, a.nbytes
        self.p = ctypes.c_void_p()
        assert H.hipMalloc(ctypes.byref(self.p), a.nbytes) == 0
        self.put(a)

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert self.H.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, shape, dtype=np.float64):
        out = np.empty(shape, dtype)
        assert self.H.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.H.hipFree(self.p)


def _states(g, seed):
    """Two nearby states of the fixture's grid: A (COLD) and B (C)."""
    a = np.array(g["final"], dtype=np.float64)
    rng = np.random.default_rng(seed)
    b = a * (1.0 + 1e-3 * rng.standard_normal(a.shape))
    return a, b


def _context_monitor(g, a, b):
    with tdt.Solver(g.mesh, g.flow, "fp64", "input") as s:
        s.set_state(b, 1e-5)
        for m in range(5):
            s.set_array(f"COLD{m + 1}", a[m])
        l2, linf = s.monitor()
        bad = False
        cold_after = np.stack([s.get_array(f"COLD{m + 1}") for m in range(5)])
    return np.asarray(l2), np.asarray(linf), bad, cold_after


def _noalloc_monitor(H, a, b, ldc_pad=0):
    n = a.shape[1]
    ldc = n + ldc_pad
    cpad = np.zeros((5, ldc))
    cpad[:, :n] = b
    dc, dcold, dout = DevBuf(H, cpad), DevBuf(H, a), DevBuf(H, np.full(11, -1.0))
    try:
        L = tlib.load()
        rc = L.tdt_monitor_noalloc(dc.p, ldc, dcold.p, n, 1, dout.p, None)
        assert rc == 0, L.tdt_last_error(None).decode()
        out = dout.get(11)
        cold_after = dcold.get((5, n))
    finally:
        for x in (dc, dcold, dout):
            x.free()
    return out[:5], out[5:10], out[10], cold_after


@pytest.mark.parametrize("pair", [("case1", "t2d_chan")])
def test_noalloc_monitor_matches_tdt_monitor(golden, pair):
    H = _hip()
    for k, name in enumerate(pair):          # the larger grid first, then the smaller
        g = golden(name)
        a, b = _states(g, seed=k)
        l2, linf, bad, cold = _context_monitor(g, a, b)
        l2n, linfn, badn, coldn = _noalloc_monitor(H, a, b, ldc_pad=37 * k)
        assert np.array_equal(l2n, l2), (name, l2n, l2)
        assert np.array_equal(linfn, linf), (name, linfn, linf)
        assert badn == 0.0 and not bad
        assert np.array_equal(coldn, b) and np.array_equal(cold, b)
        # and against the definition (numpy, another summation order)
        d = b - a
        assert np.allclose(l2n, np.sqrt((d * d).sum(axis=1) / a.shape[1]), rtol=1e-13, atol=0)
        assert np.array_equal(linfn, np.abs(d).max(axis=1))


def test_noalloc_monitor_flags_nonfinite(golden):
    H = _hip()
    g = golden("case1")
    a, b = _states(g, seed=3)
    b[3, 17] = np.nan
    _, _, badn, _ = _noalloc_monitor(H, a, b)
    assert badn == 1.0


def test_noalloc_monitor_rejects_oversize():
    L = tlib.load()
    rc = L.tdt_monitor_noalloc(ctypes.c_void_p(16), 1 << 25, ctypes.c_void_p(16), (1 << 24) + 1,
                               1, ctypes.c_void_p(16), None)
    assert rc == tlib.TDT_ERR_ARG
    assert "TDT_MONITOR_NOALLOC_MAX_CELLS" in L.tdt_last_error(None).decode()
