"""TEST INFRASTRUCTURE ONLY: ctypes loader for the C oracle (liboracle.so).

Only the tests, the smoke check and the CPU-baseline leg of the benchmark
import this module, and only as the checker.  The product never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_lib = None


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int32] * 3 + [_dp, _dp] + [_ip] * 6 + [_dp, _ip]
        L.orc_error.restype = ctypes.c_char_p
        for fn in ("orc_destroy", "orc_cold_start", "orc_timestep", "orc_dissipation",
                   "orc_primitives", "orc_apply_boundary", "orc_viscous", "orc_source_terms",
                   "orc_residual"):
            getattr(L, fn).argtypes = [ctypes.c_void_p]
            getattr(L, fn).restype = None
        L.orc_set_state.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double]
        L.orc_update.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_monitor.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_iterate.argtypes = [ctypes.c_void_p, ctypes.c_int32, _dp, _dp]
        L.orc_array.restype = _dp
        L.orc_array.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _ip]
        L.orc_scalar.restype = ctypes.c_double
        L.orc_scalar.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


class Oracle:
    """The reference's per-iteration path on the CPU, state kept reference-style."""

    def __init__(self, mesh, flow):
        L = lib()
        self._keep = [np.ascontiguousarray(a) for a in (
            mesh.x_mm, mesh.y_mm, mesh.nod, mesh.nc, mesh.parent, mesh.ghost, mesh.btype, mesh.side)]
        dp = np.array([flow.gama, flow.amol, flow.cfl, flow.prandtl, flow.prandtlt, flow.p0,
                       flow.t0, flow.pamb, flow.amachinit, flow.rtred, flow.pinf, flow.aminf,
                       flow.deltabl, flow.tflame, flow.ancomb, flow.acomb, flow.rhoprop],
                      dtype=np.float64)
        ip = np.array([flow.local, flow.npseudo, flow.iaxis], dtype=np.int32)
        k = self._keep
        self.h = L.orc_create(mesh.nods, mesh.neles, mesh.nghosts, _d(k[0]), _d(k[1]), _i(k[2]),
                              _i(k[3]), _i(k[4]), _i(k[5]), _i(k[6]), _i(k[7]), _d(dp), _i(ip))
        if not self.h:
            raise ValueError(L.orc_error().decode())
        self.neles = mesh.neles
        self.L = L

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def cold_start(self):
        self.L.orc_cold_start(self.h)

    def set_state(self, c, rinit):
        c = np.ascontiguousarray(c, dtype=np.float64)
        self.L.orc_set_state(self.h, _d(c), float(rinit))

    def timestep(self):
        self.L.orc_timestep(self.h)

    def dissipation(self):
        self.L.orc_dissipation(self.h)

    def primitives(self):
        self.L.orc_primitives(self.h)

    def apply_boundary(self):
        self.L.orc_apply_boundary(self.h)

    def viscous(self):
        self.L.orc_viscous(self.h)

    def source_terms(self):
        self.L.orc_source_terms(self.h)

    def residual(self):
        self.L.orc_residual(self.h)

    def update(self, alpha):
        self.L.orc_update(self.h, float(alpha))

    def monitor(self):
        l2 = np.zeros(5)
        linf = np.zeros(5)
        self.L.orc_monitor(self.h, _d(l2), _d(linf))
        return l2, linf

    def iterate(self, niter):
        l2 = np.zeros((niter, 5))
        linf = np.zeros((niter, 5))
        self.L.orc_iterate(self.h, int(niter), _d(l2), _d(linf))
        return l2, linf

    def array(self, name) -> np.ndarray:
        n = ctypes.c_int32(0)
        p = self.L.orc_array(self.h, name.encode(), ctypes.byref(n))
        if not p:
            raise KeyError(name)
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy()

    def scalar(self, name) -> float:
        return float(self.L.orc_scalar(self.h, name.encode()))

    def state(self) -> np.ndarray:
        return np.stack([self.array(f"C{m}")[:self.neles] for m in range(1, 6)])
