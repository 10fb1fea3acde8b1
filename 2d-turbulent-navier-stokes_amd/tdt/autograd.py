"""torch.autograd bindings of the residual and of whole iterations, with their
adjoints as the backward passes.

`residual(q, p, solver, xy)` returns R(q) for the state q (float64 tensor
[5, neles]: C1..C5 in grid.in cell order) with p = (P0, T0, AMACHINIT) as the
flow parameters and xy ([2, nods], grid.in millimetres; optional) as the node
coordinates of `solver` (a tdt.Solver, whose connectivity and other flow.in
values are used).  R is COMPUTE_RESIDUAL's R1..R5 after the primitive
variables, boundary conditions and source terms of that state
(tdt_residual_eval).  Its backward is the library's adjoint
(tdt_residual_vjp): v^T dR/dQ, v^T dR/dp and v^T dR/d(X, Y), computed on the GPU
by transposed stencil gathers.

`unrolled(p, solver, niter, xy)` returns Q^N ([5, neles]) after niter
iterations from the cold start of p = (P0, T0, AMACHINIT) on the node
coordinates xy (tdt_adjoint_forward); its backward is the reverse sweep
(tdt_adjoint_reverse), which recomputes each iteration's RK stages from states
kept every ~sqrt(niter) iterations.

The tensors stay on the CPU: the library holds the device buffers, so torch is
the autograd plumbing here, not a compute path.  libtdt is loaded before torch
so that both use the one HIP runtime the library was built against.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import lib

lib.load()

import torch  # noqa: E402  (after the library, see above)


def _use_params(solver, p) -> bool:
    """Sets the solver's P0, T0, AMACHINIT from p; True if they changed."""
    p0, t0, mach = (float(x) for x in p.detach().cpu().reshape(3))
    f = solver.flow
    if (f.p0, f.t0, f.amachinit) != (p0, t0, mach):
        solver.set_flow(dataclasses.replace(f, p0=p0, t0=t0, amachinit=mach))
        return True
    return False


def _use_nodes(solver, xy) -> bool:
    """Sets the solver's node coordinates from xy; True if they changed."""
    a = xy.detach().cpu().numpy().astype(np.float64).reshape(2, -1)
    if not np.array_equal(a, solver.xy_mm):
        solver.set_nodes(a[0], a[1])
        return True
    return False


class Residual(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q, p, xy, solver):
        _use_params(solver, p)
        _use_nodes(solver, xy)
        r = solver.residual_of(q.detach().cpu().numpy().astype(np.float64))
        ctx.solver = solver
        ctx.save_for_backward(q, p, xy)
        return torch.from_numpy(r).to(dtype=q.dtype)

    @staticmethod
    def backward(ctx, grad_r):
        q, p, xy = ctx.saved_tensors
        solver = ctx.solver
        _use_params(solver, p)
        _use_nodes(solver, xy)
        gq, gp, gxy = solver.residual_vjp(q.detach().cpu().numpy().astype(np.float64),
                                          grad_r.detach().cpu().numpy().astype(np.float64),
                                          nodes=True)
        return (torch.from_numpy(gq).to(dtype=q.dtype), torch.from_numpy(gp).to(dtype=p.dtype),
                torch.from_numpy(gxy).to(dtype=xy.dtype), None)


def residual(q, p, solver, xy=None):
    """R(q; X, Y, P0, T0, AMACHINIT) with the adjoint as its backward; xy defaults
    to the solver's current node coordinates."""
    if xy is None:
        xy = torch.from_numpy(solver.xy_mm.copy())
    return Residual.apply(q, p, xy, solver)


_runs = [0]   # labels of forward runs, so a backward can tell whether its run is the last


class Unrolled(torch.autograd.Function):
    @staticmethod
    def forward(ctx, p, xy, solver, niter):
        _use_params(solver, p)
        _use_nodes(solver, xy)
        qn = solver.adjoint_forward(int(niter))
        _runs[0] += 1
        ctx.run = _runs[0]
        solver._adjoint_run = ctx.run
        ctx.solver = solver
        ctx.niter = int(niter)
        ctx.save_for_backward(p, xy)
        return torch.from_numpy(qn).to(dtype=p.dtype)

    @staticmethod
    def backward(ctx, grad_qn):
        p, xy = ctx.saved_tensors
        solver = ctx.solver
        changed = _use_params(solver, p)
        changed = _use_nodes(solver, xy) or changed   # either discards the kept run
        if changed or getattr(solver, "_adjoint_run", None) != ctx.run:   # not this run's states
            solver.adjoint_forward(ctx.niter)
            solver._adjoint_run = ctx.run
        _, gp, gxy = solver.adjoint_reverse(grad_qn.detach().cpu().numpy().astype(np.float64),
                                            nodes=True)
        return (torch.from_numpy(gp[:3].copy()).to(dtype=p.dtype),
                torch.from_numpy(gxy).to(dtype=xy.dtype), None, None)


def unrolled(p, solver, niter, xy=None):
    """Q^N after niter iterations from the cold start of p = (P0, T0, AMACHINIT)
    on the node coordinates xy (default: the solver's), with the reverse sweep
    as its backward."""
    if xy is None:
        xy = torch.from_numpy(solver.xy_mm.copy())
    return Unrolled.apply(p, xy, solver, niter)
