"""Quantities for comparing two solutions (e.g. FP32-mixed against FP64):
wall shear stress at no-slip faces and relative L2 differences of fields."""
from __future__ import annotations

import numpy as np

from .caseio import Flow, Mesh
from .lib import primitives_from_state

SUTH = 0.00000011848     # laminar viscosity coefficient (constants.f90)


def wall_faces(mesh: Mesh, iaxis: int, min_y_mm: float = 1e-6):
    """(parent cell index, node a, node b) of the no-slip type-5 faces.  With
    IAXIS=1 a type-5 face whose midpoint is on the axis (y <= min_y_mm) is the
    axis, not a wall, and is left out; planar walls at y = 0 are kept."""
    out = []
    for k in range(mesh.nghosts):
        if mesh.btype[k] != 5:
            continue
        p, s = int(mesh.parent[k]) - 1, int(mesh.side[k]) - 1
        a, b = int(mesh.nod[s, p]) - 1, int(mesh.nod[(s + 1) % 4, p]) - 1
        if not iaxis or 0.5 * (mesh.y_mm[a] + mesh.y_mm[b]) > min_y_mm:
            out.append((p, a, b))
    return np.array(out, dtype=np.int64).reshape(-1, 3)


def wall_shear(mesh: Mesh, flow: Flow, c: np.ndarray, faces=None) -> np.ndarray:
    """Laminar wall shear stress tau_w = mu u_t / d (Pa) at no-slip faces: u_t is
    the parent cell's velocity along the face, d the distance from the cell
    centroid to the face line, mu = SUTH sqrt(AMOL) T^0.6 at the cell."""
    faces = wall_faces(mesh, flow.iaxis) if faces is None else faces
    pr = primitives_from_state(c, flow)
    x, y = mesh.x_mm * 1e-3, mesh.y_mm * 1e-3
    p, a, b = faces[:, 0], faces[:, 1], faces[:, 2]
    xc = x[mesh.nod[:, p] - 1].mean(axis=0)
    yc = y[mesh.nod[:, p] - 1].mean(axis=0)
    tx, ty = x[b] - x[a], y[b] - y[a]
    ln = np.hypot(tx, ty)
    tx, ty = tx / ln, ty / ln
    d = np.abs(tx * (yc - y[a]) - ty * (xc - x[a]))
    ut = pr["u"][p] * tx + pr["v"][p] * ty
    mu = SUTH * np.sqrt(flow.amol) * pr["T"][p] ** 0.6
    return mu * ut / d


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    """||a - b|| / ||b||."""
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def compare(mesh: Mesh, flow: Flow, c: np.ndarray, c_ref: np.ndarray) -> dict:
    """Relative L2 differences of p, rho, u, nu~ and the wall shear stress."""
    a, b = primitives_from_state(c, flow), primitives_from_state(c_ref, flow)
    out = {k: rel_l2(a[k], b[k]) for k in ("p", "rho", "u", "nut")}
    faces = wall_faces(mesh, flow.iaxis)
    if len(faces):
        out["wall_shear"] = rel_l2(wall_shear(mesh, flow, c, faces),
                                   wall_shear(mesh, flow, c_ref, faces))
    return out
