"""Objectives of a duct design study and their explicit gradients.

J = F_x + mdot, per radian of the axisymmetric duct (or per metre of depth on
a planar grid), from a state c ([5, neles], grid.in order):
  F_x   = sum over the no-slip wall faces of p_i SC_x: the axial pressure force
          on the wall, with p_i the parent cell's pressure (a type-5 ghost has
          its parent's pressure, boundary.f90:108-121) and SC_x the face's
          outward surface vector as geometry.f90:77-99 forms it (Delta y times
          the mean radius when IAXIS = 1);
  mdot  = sum over the outlet faces of (rho u)_i SC_x.
p = (GAMA - 1) (C4 - (C2^2 + C3^2) / (2 C1)), the pressure of
COMPUTE_PRIMITIVE_VARS (turbnstwod.f90:204-227) in conservative form.

duct_objective returns J, dJ/dc and the explicit dJ/d(X, Y) (per grid.in
millimetre) through the surface vectors; the path through the flow field comes
from the iteration adjoint (Solver.adjoint_reverse).  bump_node_derivatives
gives dY/da_k of the duct generator's nodes for its Hicks-Henne amplitudes.
"""
from __future__ import annotations

import numpy as np

from . import analysis, gridgen
from .caseio import Flow, Mesh


def outlet_faces(mesh: Mesh, side: int = 2):
    """(parent cell, node a, node b) of the extrapolated (type 3) faces on the
    given side (2: the right, x = max)."""
    out = []
    for k in range(mesh.nghosts):
        if mesh.btype[k] == 3 and mesh.side[k] == side:
            p, s = int(mesh.parent[k]) - 1, int(mesh.side[k]) - 1
            out.append((p, int(mesh.nod[s, p]) - 1, int(mesh.nod[(s + 1) % 4, p]) - 1))
    return np.array(out, dtype=np.int64).reshape(-1, 3)


def _scx(mesh: Mesh, iaxis: int, a, b):
    """SC_x of faces a -> b (m^2 per radian when IAXIS = 1) and its derivatives
    with respect to y_a, y_b in millimetres."""
    ya, yb = mesh.y_mm[a] * 1e-3, mesh.y_mm[b] * 1e-3
    if iaxis == 1:
        return (yb - ya) * (ya + yb) * 0.5, -ya * 1e-3, yb * 1e-3
    return yb - ya, -np.full_like(ya, 1e-3), np.full_like(yb, 1e-3)


def duct_objective(mesh: Mesh, flow: Flow, c: np.ndarray):
    """(J, dJ/dc [5, neles], explicit dJ/d(X, Y) [2, nods] per mm)."""
    c = np.asarray(c, dtype=np.float64)
    g1 = flow.gama - 1.0
    dc = np.zeros_like(c)
    dxy = np.zeros((2, mesh.nods))
    wall = analysis.wall_faces(mesh, flow.iaxis)
    p_, a, b = wall[:, 0], wall[:, 1], wall[:, 2]
    sx, dya, dyb = _scx(mesh, flow.iaxis, a, b)
    c1, c2, c3, c4 = c[0, p_], c[1, p_], c[2, p_], c[3, p_]
    p = g1 * (c4 - 0.5 * (c2 * c2 + c3 * c3) / c1)
    fx = float(np.sum(p * sx))
    np.add.at(dc[0], p_, g1 * 0.5 * (c2 * c2 + c3 * c3) / (c1 * c1) * sx)
    np.add.at(dc[1], p_, -g1 * c2 / c1 * sx)
    np.add.at(dc[2], p_, -g1 * c3 / c1 * sx)
    np.add.at(dc[3], p_, g1 * sx)
    np.add.at(dxy[1], a, p * dya)
    np.add.at(dxy[1], b, p * dyb)
    out = outlet_faces(mesh)
    q_, a, b = out[:, 0], out[:, 1], out[:, 2]
    sx, dya, dyb = _scx(mesh, flow.iaxis, a, b)
    mdot = float(np.sum(c[1, q_] * sx))
    np.add.at(dc[1], q_, sx)
    np.add.at(dxy[1], a, c[1, q_] * dya)
    np.add.at(dxy[1], b, c[1, q_] * dyb)
    return fx + mdot, dc, dxy


def bump_node_derivatives(nbumps: int = 8, h: float = 1e-3, **duct_kw):
    """dX/da_k, dY/da_k [nbumps, 2, nods] of gridgen.duct's nodes at zero
    amplitudes, by central differences of the generator in each amplitude
    (step h mm; the nodes are smooth in the amplitudes and rounded to 1e-10 mm,
    so the differences are good to ~1e-7)."""
    out = []
    for k in range(nbumps):
        amps = np.zeros(nbumps)
        amps[k] = h
        mp = gridgen.duct(bumps=list(amps), **duct_kw)
        mm = gridgen.duct(bumps=list(-amps), **duct_kw)
        out.append(np.stack([(mp.x_mm - mm.x_mm) / (2 * h), (mp.y_mm - mm.y_mm) / (2 * h)]))
    return np.stack(out)
