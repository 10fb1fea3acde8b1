"""Readers and writers for the solver's case files.

The formats are those of the Fortran solver:
  * flow.in    one list-directed record of 17 fields (turbnstwod.f90:20-23)
  * grid.in    "NODS NELES", NODS node records "N X Y [Z]" (mm), then NELES
               element records "I N1 N2 N3 N4 NC1 NC2 NC3 NC4" (geometry.f90:31-44)
  * bc.in      "NGHOSTS", then NGHOSTS records "PARENT GHOST TYPE SIDE"
               (geometry.f90:47-52)
  * restart.in NELES records "C1 C2 C3 C4 C5", then RTRED, NPSEUDO, RINIT,
               NITEROLD, one per record (turbnstwod.f90:73-85, 186-195)

All ids stay 1-based as in the files.  Connectivity arrays are column-major
[4, NELES] (NOD(I,K) -> nod[K-1, I-1]), the layout the C ABI takes.
"""
from __future__ import annotations

import dataclasses
import os
import re
from typing import Optional

import numpy as np

FLOW_FIELDS = ("fnamegrid", "fnamebc", "gama", "amol", "local", "niter", "istart",
               "cfl", "prandtl", "prandtlt", "p0", "t0", "pamb", "amachinit",
               "rtred", "npseudo", "iaxis")
_INT_FIELDS = {"local", "niter", "istart", "npseudo", "iaxis"}


@dataclasses.dataclass
class Flow:
    """The 17 flow.in fields plus the inputs of BC types 7 and 6.

    pinf/aminf/deltabl (BC 7, boundary.f90:312-317) and tflame/ancomb/acomb/rhoprop
    (BC 6: flame temperature, burn-rate exponent and coefficient, propellant
    density; boundary.f90:138-153) are never assigned by the reference, which
    reads them as zero; they default to 0 here and are set only by callers that
    choose to define those boundary types.
    """
    fnamegrid: str = "grid.in"
    fnamebc: str = "bc.in"
    gama: float = 1.4
    amol: float = 28.96
    local: int = 1
    niter: int = 100
    istart: int = 0
    cfl: float = 0.25
    prandtl: float = 0.72
    prandtlt: float = 0.9
    p0: float = 1.2e5
    t0: float = 298.0
    pamb: float = -1.0
    amachinit: float = 0.3
    rtred: float = 1e-5
    npseudo: int = 1
    iaxis: int = 0
    pinf: float = 0.0
    aminf: float = 0.0
    deltabl: float = 0.0
    tflame: float = 0.0
    ancomb: float = 0.0
    acomb: float = 0.0
    rhoprop: float = 0.0


@dataclasses.dataclass
class Mesh:
    """Grid and boundary lists exactly as read from grid.in / bc.in."""
    x_mm: np.ndarray          # [nods] float64
    y_mm: np.ndarray          # [nods] float64
    nod: np.ndarray           # [4, neles] int32, 1-based node ids
    nc: np.ndarray            # [4, neles] int32, 1-based cell/ghost ids
    parent: np.ndarray        # [nghosts] int32
    ghost: np.ndarray         # [nghosts] int32
    btype: np.ndarray         # [nghosts] int32
    side: np.ndarray          # [nghosts] int32

    @property
    def nods(self) -> int:
        return int(self.x_mm.shape[0])

    @property
    def neles(self) -> int:
        return int(self.nod.shape[1])

    @property
    def nghosts(self) -> int:
        return int(self.parent.shape[0])

    def arrays(self) -> dict:
        return {"x_mm": self.x_mm, "y_mm": self.y_mm, "nod": self.nod, "nc": self.nc,
                "parent": self.parent, "ghost": self.ghost, "btype": self.btype,
                "side": self.side}

    def digest(self) -> str:
        """SHA-256 over the mesh arrays (fixed key order, native byte layout): names
        a generated grid in a fixture without storing it."""
        import hashlib
        h = hashlib.sha256()
        for k, v in self.arrays().items():
            h.update(k.encode())
            h.update(np.ascontiguousarray(v).tobytes())
        return h.hexdigest()

    @staticmethod
    def from_arrays(d) -> "Mesh":
        return Mesh(np.ascontiguousarray(d["x_mm"], dtype=np.float64),
                    np.ascontiguousarray(d["y_mm"], dtype=np.float64),
                    np.ascontiguousarray(d["nod"], dtype=np.int32),
                    np.ascontiguousarray(d["nc"], dtype=np.int32),
                    np.ascontiguousarray(d["parent"], dtype=np.int32),
                    np.ascontiguousarray(d["ghost"], dtype=np.int32),
                    np.ascontiguousarray(d["btype"], dtype=np.int32),
                    np.ascontiguousarray(d["side"], dtype=np.int32))


def _fortran_float(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "e"))


def _list_directed_tokens(text: str):
    # List-directed input: blanks, tabs, commas and newlines separate values;
    # a slash ends the record.  Quoted strings are not used by the case files.
    text = text.split("/")[0] if "/" in text and "'" not in text and '"' not in text else text
    return [t for t in re.split(r"[\s,]+", text) if t]


def read_flow(path: str) -> Flow:
    with open(path) as fh:
        toks = _list_directed_tokens(fh.read())
    if len(toks) < 17:
        raise ValueError(f"{path}: expected 17 fields, found {len(toks)}")
    kw = {}
    for name, tok in zip(FLOW_FIELDS, toks[:17]):
        if name in ("fnamegrid", "fnamebc"):
            kw[name] = tok.strip("'\"")
        elif name in _INT_FIELDS:
            kw[name] = int(tok)
        else:
            kw[name] = _fortran_float(tok)
    return Flow(**kw)


def write_flow(path: str, f: Flow) -> None:
    vals = [f.fnamegrid, f.fnamebc, repr(float(f.gama)), repr(float(f.amol)), str(int(f.local)),
            str(int(f.niter)), str(int(f.istart)), repr(float(f.cfl)), repr(float(f.prandtl)),
            repr(float(f.prandtlt)), repr(float(f.p0)), repr(float(f.t0)), repr(float(f.pamb)),
            repr(float(f.amachinit)), repr(float(f.rtred)), str(int(f.npseudo)), str(int(f.iaxis))]
    with open(path, "w") as fh:
        fh.write("\t".join(vals) + "\n\n" + ", ".join(n.upper() for n in FLOW_FIELDS) + "\n")


def read_grid_bc(grid_path: str, bc_path: str) -> Mesh:
    with open(grid_path) as fh:
        head = fh.readline().split()
        nods, neles = int(head[0]), int(head[1])
        x = np.zeros(nods)
        y = np.zeros(nods)
        for _ in range(nods):
            t = fh.readline().split()
            n = int(t[0])
            x[n - 1] = _fortran_float(t[1])
            y[n - 1] = _fortran_float(t[2])
        el = np.loadtxt(fh, dtype=np.int64, max_rows=neles, ndmin=2)
    if el.shape != (neles, 9):
        raise ValueError(f"{grid_path}: expected {neles} element records of 9 integers")
    # READ(21,*) N, NOD(I,1..4), NC1..4(I): record I goes to slot I regardless of N
    nod = np.ascontiguousarray(el[:, 1:5].T, dtype=np.int32)
    nc = np.ascontiguousarray(el[:, 5:9].T, dtype=np.int32)
    with open(bc_path) as fh:
        ng = int(fh.readline().split()[0])
        bc = np.loadtxt(fh, dtype=np.int64, max_rows=ng, ndmin=2) if ng else np.zeros((0, 4), np.int64)
    bc = bc.reshape(-1, 4)
    return Mesh(x, y, nod, nc, bc[:, 0].astype(np.int32), bc[:, 1].astype(np.int32),
                bc[:, 2].astype(np.int32), bc[:, 3].astype(np.int32))


def write_grid_bc(grid_path: str, bc_path: str, m: Mesh) -> None:
    lines = [f"{m.nods:12d}{m.neles:12d}"]
    for i in range(m.nods):
        lines.append(f"{i + 1:8d} {m.x_mm[i]:18.10f} {m.y_mm[i]:18.10f} {0.0:18.10f}")
    nod, nc = m.nod, m.nc
    for i in range(m.neles):
        lines.append("%8d%8d%8d%8d%8d%8d%8d%8d%8d" % (i + 1, nod[0, i], nod[1, i], nod[2, i], nod[3, i],
                                                      nc[0, i], nc[1, i], nc[2, i], nc[3, i]))
    with open(grid_path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    bl = [f"{m.nghosts:12d}"]
    for g in range(m.nghosts):
        bl.append(f"{m.parent[g]:12d}{m.ghost[g]:12d}{m.btype[g]:12d}{m.side[g]:12d}")
    with open(bc_path, "w") as fh:
        fh.write("\n".join(bl) + "\n")


def read_restart(path: str, neles: int):
    """Return (c[5, neles], rtred, npseudo, rinit, niterold)."""
    with open(path) as fh:
        toks = _list_directed_tokens(fh.read())
    need = 5 * neles + 4
    if len(toks) < need:
        raise ValueError(f"{path}: truncated restart file ({len(toks)} of {need} values)")
    c = np.array([_fortran_float(t) for t in toks[:5 * neles]]).reshape(neles, 5).T.copy()
    rest = toks[5 * neles:5 * neles + 4]
    return c, _fortran_float(rest[0]), int(rest[1]), _fortran_float(rest[2]), int(rest[3])


def write_restart(path: str, c: np.ndarray, rtred: float, npseudo: int, rinit: float,
                  niterold: int) -> None:
    """Writes values with shortest round-trip digits so a read returns the same bits."""
    out = []
    for i in range(c.shape[1]):
        out.append(" " + " ".join(repr(float(v)) for v in c[:, i]))
    out += [" " + repr(float(rtred)), " " + str(int(npseudo)), " " + repr(float(rinit)),
            " " + str(int(niterold))]
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def load_case(case_dir: str) -> tuple:
    flow = read_flow(os.path.join(case_dir, "flow.in"))
    mesh = read_grid_bc(os.path.join(case_dir, flow.fnamegrid), os.path.join(case_dir, flow.fnamebc))
    return flow, mesh


def write_case(case_dir: str, flow: Flow, mesh: Mesh, restart: Optional[tuple] = None) -> None:
    os.makedirs(case_dir, exist_ok=True)
    write_flow(os.path.join(case_dir, "flow.in"), flow)
    write_grid_bc(os.path.join(case_dir, flow.fnamegrid), os.path.join(case_dir, flow.fnamebc), mesh)
    if restart is not None:
        write_restart(os.path.join(case_dir, "restart.in"), *restart)
