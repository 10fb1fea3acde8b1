"""ctypes binding of libtdt (include/tdt.h).

`Solver` mirrors the reference's routine boundary: one method per Fortran
subroutine of the per-iteration path (timestep, dissipation, primitives,
apply_boundary, viscous, source_terms, residual, update, monitor) plus the
fused `iterate`.  All arrays are numpy, in the caller's original ordering.

The library is the only compute path: if lib/libtdt.so is missing or no GPU
is available, construction raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

from .caseio import Flow, Mesh

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# TDT_LIBRARY selects another build of the library (e.g. to compare two builds)
LIB_PATH = os.environ.get("TDT_LIBRARY") or os.path.join(PKG_DIR, "lib", "libtdt.so")
HEADER_PATH = os.path.join(os.path.dirname(PKG_DIR), "include", "tdt.h")

TDT_OK, TDT_ERR_ARG, TDT_ERR_MESH, TDT_ERR_DEVICE, TDT_ERR_NONFINITE, TDT_ERR_UNSUPPORTED = (
    0, -1, -2, -3, -4, -5)
PRECISION = {"fp64": 0, "fp32-mixed": 1}
ORDERING = {"input": 0, "rcm": 1, "auto": 2, "boundary-last": 3}
ERROR_NAMES = {-1: "TDT_ERR_ARG", -2: "TDT_ERR_MESH", -3: "TDT_ERR_DEVICE",
               -4: "TDT_ERR_NONFINITE", -5: "TDT_ERR_UNSUPPORTED"}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


class TdtError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


class _Mesh(ctypes.Structure):
    _fields_ = [("nods", ctypes.c_int32), ("neles", ctypes.c_int32), ("nghosts", ctypes.c_int32),
                ("x_mm", _dp), ("y_mm", _dp), ("nod", _ip), ("nc", _ip), ("parent", _ip),
                ("ghost", _ip), ("type", _ip), ("side", _ip)]


class _Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "gama", "amol", "cfl", "prandtl", "prandtlt", "p0", "t0", "pamb", "amachinit", "rtred",
        "pinf", "aminf", "deltabl", "tflame", "ancomb", "acomb", "rhoprop")] + [
            (n, ctypes.c_int32) for n in (
            "local", "npseudo", "iaxis", "reserved")]


_lib = None


def load(path: str = LIB_PATH):
    """Loads libtdt.so and declares the prototypes of include/tdt.h."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with `make -C {PKG_DIR}` (there is no CPU fallback)")
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    L.tdt_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(_Mesh), ctypes.POINTER(_Params),
                             ctypes.c_int32, ctypes.c_int32]
    L.tdt_last_error.restype = ctypes.c_char_p
    L.tdt_last_error.argtypes = [vp]
    L.tdt_version.restype = ctypes.c_char_p
    L.tdt_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    L.tdt_cold_start.argtypes = [vp, _dp]
    L.tdt_set_state.argtypes = [vp, _dp, ctypes.c_double]
    L.tdt_get_state.argtypes = [vp, _dp]
    for fn in ("tdt_destroy", "tdt_timestep", "tdt_dissipation", "tdt_primitives",
               "tdt_apply_boundary", "tdt_viscous", "tdt_source_terms", "tdt_residual",
               "tdt_synchronize"):
        getattr(L, fn).argtypes = [vp]
    L.tdt_update.argtypes = [vp, ctypes.c_double]
    L.tdt_monitor.argtypes = [vp, _dp, _dp]
    L.tdt_iterate.argtypes = [vp, ctypes.c_int32, _dp, _dp, _ip]
    L.tdt_get_array.argtypes = [vp, ctypes.c_char_p, _dp, ctypes.c_int32]
    L.tdt_set_array.argtypes = [vp, ctypes.c_char_p, _dp, ctypes.c_int32]
    L.tdt_set_profiling.argtypes = [vp, ctypes.c_int32]
    L.tdt_kernel_times.argtypes = [vp, _dp]
    L.tdt_kernel_set.argtypes = [vp, _ip]
    L.tdt_postprocess.argtypes = [vp, _dp]
    L.tdt_residual_eval.argtypes = [vp, _dp, _dp]
    L.tdt_residual_vjp.argtypes = [vp, _dp, _dp, _dp, _dp, _dp]
    L.tdt_set_nodes.argtypes = [vp, _dp, _dp]
    L.tdt_set_params.argtypes = [vp, ctypes.POINTER(_Params)]
    L.tdt_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int32, ctypes.POINTER(_Mesh),
                                   ctypes.POINTER(_Params), ctypes.c_int32, ctypes.c_int32]
    L.tdt_batch_cold_start.argtypes = [vp, _dp]
    L.tdt_batch_set_state.argtypes = [vp, ctypes.c_int32, _dp, ctypes.c_double]
    L.tdt_batch_get_state.argtypes = [vp, ctypes.c_int32, _dp]
    L.tdt_batch_iterate.argtypes = [vp, ctypes.c_int32, _dp, _dp, _ip]
    L.tdt_operators_eval.argtypes = [vp, _dp, _dp]
    L.tdt_operators_vjp.argtypes = [vp, _dp, ctypes.c_double, _dp, _dp, _dp, _dp]
    L.tdt_iteration_vjp.argtypes = [vp, _dp, ctypes.c_double, _dp, _dp, _dp, _dp]
    L.tdt_adjoint_fixed_point.argtypes = [vp, _dp, ctypes.c_double, _dp, ctypes.c_int32,
                                          ctypes.c_double, _dp, _dp, _dp, _ip, _dp]
    L.tdt_adjoint_forward.argtypes = [vp, ctypes.c_int32, _dp, ctypes.c_double, _dp]
    L.tdt_adjoint_reverse.argtypes = [vp, _dp, _dp, _dp, _dp]
    # device pointers and a hipStream_t, passed as addresses
    L.tdt_monitor_noalloc.argtypes = [vp, ctypes.c_int64, vp, ctypes.c_int32, ctypes.c_int32, vp,
                                      vp]
    _lib = L
    return L


def header_symbols(path: str = HEADER_PATH):
    """Function names declared in include/tdt.h."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tdt_[a-z0-9_]+)\s*\(", text)))


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def make_params(flow: Flow) -> _Params:
    return _Params(flow.gama, flow.amol, flow.cfl, flow.prandtl, flow.prandtlt, flow.p0, flow.t0,
                   flow.pamb, flow.amachinit, flow.rtred, flow.pinf, flow.aminf, flow.deltabl,
                   flow.tflame, flow.ancomb, flow.acomb, flow.rhoprop,
                   int(flow.local), int(flow.npseudo), int(flow.iaxis), 0)


def _mesh_struct(mesh: Mesh, keep: list) -> _Mesh:
    """The tdt_mesh view of a Mesh; the arrays it points to are appended to keep."""
    k = [np.ascontiguousarray(a) for a in (
        mesh.x_mm.astype(np.float64), mesh.y_mm.astype(np.float64), mesh.nod.astype(np.int32),
        mesh.nc.astype(np.int32), mesh.parent.astype(np.int32), mesh.ghost.astype(np.int32),
        mesh.btype.astype(np.int32), mesh.side.astype(np.int32))]
    keep.extend(k)
    return _Mesh(mesh.nods, mesh.neles, mesh.nghosts, _d(k[0]), _d(k[1]), _i(k[2]), _i(k[3]),
                 _i(k[4]), _i(k[5]), _i(k[6]), _i(k[7]))


class Solver:
    """One device-resident case: mesh, geometry, state Q and work arrays."""

    CELL_GHOST = {"U", "V", "T", "P", "DT", "ANU", "ANURT", "XCC", "YCC", "VOL"} | \
        {f"{p}{m}" for p in ("C", "F", "G", "D2Q") for m in range(1, 6)}
    NODE = {"UN", "VN", "TN", "RHON", "RN"}
    SCALAR = {"DTMIN", "RTOL", "RINIT"}

    def __init__(self, mesh: Mesh, flow: Flow, precision: str = "fp64", ordering: str = "auto"):
        self.L = load()
        self.mesh, self.flow = mesh, flow
        self._keep = []
        m = _mesh_struct(mesh, self._keep)
        p = make_params(flow)
        h = ctypes.c_void_p()
        rc = self.L.tdt_create(ctypes.byref(h), ctypes.byref(m), ctypes.byref(p),
                               PRECISION[precision], ORDERING[ordering])
        if rc != TDT_OK:
            raise TdtError(rc, self.L.tdt_last_error(None).decode())
        self.h = h
        info = (ctypes.c_int64 * 8)()
        self.L.tdt_info(h, info)
        self.neles, self.nghosts, self.nods, self.nmax = info[0], info[1], info[2], info[3]
        self.bandwidth, self.ordering_used = info[4], info[6]
        self.precision = precision
        self.xy_mm = np.stack([mesh.x_mm, mesh.y_mm]).astype(np.float64)

    def close(self):
        if getattr(self, "h", None):
            self.L.tdt_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != TDT_OK:
            raise TdtError(rc, self.L.tdt_last_error(self.h).decode())

    # ---- state -----------------------------------------------------------------------
    def cold_start(self) -> float:
        r = ctypes.c_double(0.0)
        self._check(self.L.tdt_cold_start(self.h, ctypes.byref(r)))
        return r.value

    def set_state(self, c: np.ndarray, rinit: float) -> None:
        c = np.ascontiguousarray(c, dtype=np.float64).reshape(5, self.neles)
        self._check(self.L.tdt_set_state(self.h, _d(c), float(rinit)))

    def get_state(self) -> np.ndarray:
        out = np.empty((5, self.neles))
        self._check(self.L.tdt_get_state(self.h, _d(out)))
        return out

    # ---- routine-level entry points ---------------------------------------------------
    def timestep(self): self._check(self.L.tdt_timestep(self.h))
    def dissipation(self): self._check(self.L.tdt_dissipation(self.h))
    def primitives(self): self._check(self.L.tdt_primitives(self.h))
    def apply_boundary(self): self._check(self.L.tdt_apply_boundary(self.h))
    def viscous(self): self._check(self.L.tdt_viscous(self.h))
    def source_terms(self): self._check(self.L.tdt_source_terms(self.h))
    def residual(self): self._check(self.L.tdt_residual(self.h))
    def update(self, alpha: float): self._check(self.L.tdt_update(self.h, float(alpha)))

    def monitor(self):
        l2 = np.zeros(5)
        linf = np.zeros(5)
        self._check(self.L.tdt_monitor(self.h, _d(l2), _d(linf)))
        return l2, linf

    def postprocess(self) -> dict:
        """POSTPROCESS (solver_state.f90:230-283) on the device: node values of the
        tec_plot.dat columns U, V, P [bar], RHO, MACH, T from the current P, U, V, T
        arrays, in grid.in node order."""
        out = np.empty((6, self.nods))
        self._check(self.L.tdt_postprocess(self.h, _d(out)))
        return dict(zip(("U", "V", "P_bar", "RHO", "MACH", "T"), out))

    # ---- fused hot path ----------------------------------------------------------------
    def iterate(self, niter: int, history: bool = True, check: bool = True):
        """Returns (l2[niter,5], linf[niter,5], first_nonfinite)."""
        l2 = np.zeros((niter, 5)) if history else None
        linf = np.zeros((niter, 5)) if history else None
        bad = ctypes.c_int32(0)
        rc = self.L.tdt_iterate(self.h, int(niter), _d(l2) if history else None,
                                _d(linf) if history else None, ctypes.byref(bad))
        if rc == TDT_ERR_NONFINITE and not check:
            return l2, linf, bad.value
        self._check(rc)
        return l2, linf, bad.value

    # ---- residual and its adjoint -----------------------------------------------------
    def residual_of(self, c: np.ndarray) -> np.ndarray:
        """R(c): COMPUTE_RESIDUAL's R1..R5 after primitives, boundary conditions and
        source terms on the state c ([5, neles]); the context's state is untouched."""
        c = np.ascontiguousarray(c, dtype=np.float64).reshape(5, self.neles)
        r = np.empty((5, self.neles))
        self._check(self.L.tdt_residual_eval(self.h, _d(c), _d(r)))
        return r

    def residual_vjp(self, c: np.ndarray, v: np.ndarray, nodes: bool = False):
        """(v^T dR/dQ [5, neles], v^T dR/d(P0, T0, AMACHINIT) [3]) at the state c,
        and with nodes=True also v^T dR/d(X, Y) [2, nods] per grid.in millimetre."""
        c = np.ascontiguousarray(c, dtype=np.float64).reshape(5, self.neles)
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(5, self.neles)
        g = np.empty((5, self.neles))
        gp = np.zeros(3)
        gxy = np.empty((2, self.nods)) if nodes else None
        self._check(self.L.tdt_residual_vjp(self.h, _d(c), _d(v), _d(g), _d(gp),
                                            _d(gxy) if nodes else None))
        return (g, gp, gxy) if nodes else (g, gp)

    # ---- adjoint of whole iterations ----------------------------------------------------
    def operators_of(self, c: np.ndarray) -> dict:
        """D241..D245, VISU..VISR and DT of the state c as the iteration adjoint
        evaluates them (the routine entry points' values)."""
        c = np.ascontiguousarray(c, dtype=np.float64).reshape(5, self.neles)
        out = np.empty((10, self.neles))
        self._check(self.L.tdt_operators_eval(self.h, _d(c), _d(out)))
        names = [f"D24{m}" for m in range(1, 6)] + ["VISU", "VISV", "VIST", "VISR", "DT"]
        return dict(zip(names, out))

    def operators_vjp(self, c: np.ndarray, rinit: float, w: np.ndarray, nodes: bool = False):
        """Weights w [7, neles] on E = D + VIS (5 rows), DT and VOL at the state c:
        (w^T d(E, DT)/dQ [5, neles], w^T dE/d(P0, T0, RINIT) [3]) and with
        nodes=True also w^T d(E, DT, VOL)/d(X, Y) [2, nods] per millimetre."""
        c = np.ascontiguousarray(c, dtype=np.float64).reshape(5, self.neles)
        w = np.ascontiguousarray(w, dtype=np.float64).reshape(7, self.neles)
        g = np.empty((5, self.neles))
        gp = np.zeros(3)
        gxy = np.empty((2, self.nods)) if nodes else None
        self._check(self.L.tdt_operators_vjp(self.h, _d(c), float(rinit), _d(w), _d(g), _d(gp),
                                             _d(gxy) if nodes else None))
        return (g, gp, gxy) if nodes else (g, gp)

    def iteration_vjp(self, c: np.ndarray, rinit: float, mu: np.ndarray, nodes: bool = False):
        """One iteration Q^n = c -> Q^{n+1}: (mu^T dQ^{n+1}/dQ^n [5, neles],
        mu^T dQ^{n+1}/d(P0, T0, RINIT) [3]) and with nodes=True also
        mu^T dQ^{n+1}/d(X, Y) [2, nods] per grid.in millimetre."""
        c = np.ascontiguousarray(c, dtype=np.float64).reshape(5, self.neles)
        mu = np.ascontiguousarray(mu, dtype=np.float64).reshape(5, self.neles)
        lam = np.empty((5, self.neles))
        gp = np.zeros(3)
        gxy = np.empty((2, self.nods)) if nodes else None
        self._check(self.L.tdt_iteration_vjp(self.h, _d(c), float(rinit), _d(mu), _d(lam),
                                             _d(gp), _d(gxy) if nodes else None))
        return (lam, gp, gxy) if nodes else (lam, gp)

    def adjoint_fixed_point(self, c: np.ndarray, rinit: float, dj_dc: np.ndarray,
                            max_iter: int = 20000, tol: float = 1e-12, nodes: bool = False):
        """Adjoint at a steady state c: (lam = (I - Phi_Q^T)^-1 dJ/dc [5, neles],
        lam^T dPhi/d(P0, T0, RINIT) [3], [lam^T dPhi/d(X, Y) [2, nods] with nodes],
        iterations done, last relative change)."""
        c = np.ascontiguousarray(c, dtype=np.float64).reshape(5, self.neles)
        g = np.ascontiguousarray(dj_dc, dtype=np.float64).reshape(5, self.neles)
        lam = np.empty((5, self.neles))
        gp = np.zeros(3)
        gxy = np.empty((2, self.nods)) if nodes else None
        it = ctypes.c_int32(0)
        ch = ctypes.c_double(0.0)
        self._check(self.L.tdt_adjoint_fixed_point(self.h, _d(c), float(rinit), _d(g),
                                                   int(max_iter), float(tol), _d(lam), _d(gp),
                                                   _d(gxy) if nodes else None,
                                                   ctypes.byref(it), ctypes.byref(ch)))
        return (lam, gp, gxy, it.value, ch.value) if nodes else (lam, gp, it.value, ch.value)

    def adjoint_forward(self, niter: int, c0=None, rinit: float = 0.0) -> np.ndarray:
        """niter iterations from c0 (None: the cold start), keeping states for
        adjoint_reverse; returns Q^N [5, neles]."""
        cn = np.empty((5, self.neles))
        if c0 is None:
            self._check(self.L.tdt_adjoint_forward(self.h, int(niter), None, 0.0, _d(cn)))
        else:
            c0 = np.ascontiguousarray(c0, dtype=np.float64).reshape(5, self.neles)
            self._check(self.L.tdt_adjoint_forward(self.h, int(niter), _d(c0), float(rinit),
                                                   _d(cn)))
        return cn

    def adjoint_reverse(self, dj_dcn: np.ndarray, nodes: bool = False):
        """Reverse sweep of the last adjoint_forward: (dJ/dQ^0 [5, neles],
        dJ/d(P0, T0, AMACHINIT, RINIT) [4]) and with nodes=True also dJ/d(X, Y)
        [2, nods] per grid.in millimetre."""
        g = np.ascontiguousarray(dj_dcn, dtype=np.float64).reshape(5, self.neles)
        lam = np.empty((5, self.neles))
        gp = np.zeros(4)
        gxy = np.empty((2, self.nods)) if nodes else None
        self._check(self.L.tdt_adjoint_reverse(self.h, _d(g), _d(lam), _d(gp),
                                               _d(gxy) if nodes else None))
        return (lam, gp, gxy) if nodes else (lam, gp)

    def set_nodes(self, x_mm: np.ndarray, y_mm: np.ndarray) -> None:
        """New node coordinates (grid.in order, mm); the geometry is recomputed."""
        x = np.ascontiguousarray(x_mm, dtype=np.float64).reshape(self.nods)
        y = np.ascontiguousarray(y_mm, dtype=np.float64).reshape(self.nods)
        self._check(self.L.tdt_set_nodes(self.h, _d(x), _d(y)))
        self.xy_mm = np.stack([x, y])

    def set_flow(self, flow: Flow) -> None:
        """New flow.in values (same IAXIS) for this context."""
        p = make_params(flow)
        self._check(self.L.tdt_set_params(self.h, ctypes.byref(p)))
        self.flow = flow

    # ---- arrays ----------------------------------------------------------------------
    def array_len(self, name: str) -> int:
        if name in self.SCALAR:
            return 1
        if name in self.CELL_GHOST:
            return self.nmax
        if name in self.NODE:
            return self.nods
        return self.neles

    def get_array(self, name: str) -> np.ndarray:
        n = self.array_len(name)
        out = np.zeros(n)
        self._check(self.L.tdt_get_array(self.h, name.encode(), _d(out), n))
        return out

    def set_array(self, name: str, values: np.ndarray) -> None:
        n = self.array_len(name)
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.shape != (n,):
            raise ValueError(f"{name}: expected {n} values, got {v.shape}")
        self._check(self.L.tdt_set_array(self.h, name.encode(), _d(v), n))

    def set_profiling(self, on: bool) -> None:
        self._check(self.L.tdt_set_profiling(self.h, 1 if on else 0))

    def kernel_times(self) -> np.ndarray:
        out = np.zeros(6)
        self._check(self.L.tdt_kernel_times(self.h, _d(out)))
        return out

    KERNEL_SETS = ("merged", "split", "quad", "face")
    _UNIFORM = {0: "merged", 2: "split", 13: "quad", 6: "face"}

    def kernel_set(self) -> str:
        """The fused kernel set in use: "merged", "split", "quad", "face" (FP32-mixed:
        each viscous face evaluated once), or for a mixed
        choice e.g. "pre=quad,stage1=merged,stages=lane", with "+preload" when the
        FP64 one-lane stage kernels load all neighbour states up front and "+nodes"
        when the FP64 dissipation / viscous kernel copies its node records into LDS
        ("" before the first tdt_iterate)."""
        out = np.zeros(1, dtype=np.int32)
        self._check(self.L.tdt_kernel_set(self.h, _i(out)))
        v = int(out[0])
        if v < 0:
            return ""
        pl = ("+preload" if v & 16 else "") + ("+nodes" if v & 32 else "")
        v &= 15
        if v in self._UNIFORM:
            return self._UNIFORM[v] + pl
        st1 = {0: "merged", 2: "split", 4: "quad", 6: "face"}[v & 6]
        return (f"pre={'quad' if v & 1 else 'lane'},stage1={st1},"
                f"stages={'quad' if v & 8 else 'lane'}{pl}")

    def synchronize(self):
        self._check(self.L.tdt_synchronize(self.h))


class Batch:
    """Independent cases advanced together (tdt_batch_*): one fused iteration over
    the concatenated cells of every case, each case with its own flow constants.
    Per case, the results are those of a Solver of that case alone."""

    def __init__(self, cases, precision: str = "fp64", ordering: str = "auto"):
        """cases: a sequence of (Mesh, Flow)."""
        self.L = load()
        self.cases = list(cases)
        self._keep = []
        meshes = (_Mesh * len(self.cases))(*[_mesh_struct(m, self._keep) for m, _ in self.cases])
        params = (_Params * len(self.cases))(*[make_params(f) for _, f in self.cases])
        h = ctypes.c_void_p()
        rc = self.L.tdt_batch_create(ctypes.byref(h), len(self.cases), meshes, params,
                                     PRECISION[precision], ORDERING[ordering])
        if rc != TDT_OK:
            raise TdtError(rc, self.L.tdt_last_error(None).decode())
        self.h = h
        self.ncases = len(self.cases)
        self.neles = [m.neles for m, _ in self.cases]
        self.precision = precision

    close = Solver.close
    __del__ = Solver.__del__
    __enter__ = Solver.__enter__
    __exit__ = Solver.__exit__
    _check = Solver._check
    _UNIFORM = Solver._UNIFORM
    kernel_set = Solver.kernel_set
    set_profiling = Solver.set_profiling
    kernel_times = Solver.kernel_times
    synchronize = Solver.synchronize

    def cold_start(self) -> np.ndarray:
        r = np.zeros(self.ncases)
        self._check(self.L.tdt_batch_cold_start(self.h, _d(r)))
        return r

    def set_state(self, icase: int, c: np.ndarray, rinit: float) -> None:
        c = np.ascontiguousarray(c, dtype=np.float64).reshape(5, self.neles[icase])
        self._check(self.L.tdt_batch_set_state(self.h, int(icase), _d(c), float(rinit)))

    def get_state(self, icase: int) -> np.ndarray:
        out = np.empty((5, self.neles[icase]))
        self._check(self.L.tdt_batch_get_state(self.h, int(icase), _d(out)))
        return out

    def iterate(self, niter: int, history: bool = True, check: bool = True):
        """Returns (l2[niter, ncases, 5], linf[niter, ncases, 5], first_nonfinite)."""
        l2 = np.zeros((niter, self.ncases, 5)) if history else None
        linf = np.zeros((niter, self.ncases, 5)) if history else None
        bad = ctypes.c_int32(0)
        rc = self.L.tdt_batch_iterate(self.h, int(niter), _d(l2) if history else None,
                                      _d(linf) if history else None, ctypes.byref(bad))
        if rc == TDT_ERR_NONFINITE and not check:
            return l2, linf, bad.value
        self._check(rc)
        return l2, linf, bad.value


def version() -> str:
    return load().tdt_version().decode()


def primitives_from_state(c: np.ndarray, flow: Flow) -> dict:
    """rho, u, v, p, nu~ from C1..C5 with the reference's formulas (turbnstwod.f90:204-227)."""
    rr = 8314.0 / flow.amol
    cv = rr / (flow.gama - 1.0)
    rho = c[0]
    u = c[1] / rho
    v = c[2] / rho
    e = c[3] / rho
    t = (e - 0.5 * (u * u + v * v)) / cv
    p = rho * rr * t
    return {"rho": rho, "u": u, "v": v, "p": p, "nut": c[4] / rho, "T": t}
