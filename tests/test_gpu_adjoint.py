"""The residual and its adjoint on the GPU (tdt_residual_eval / tdt_residual_vjp).

R(Q) is COMPUTE_RESIDUAL's R1..R5 after primitives, boundary conditions and
source terms.  Its value must be the routine path's, bit for bit; its adjoint
v^T dR/dQ (the transposed stencil gather of the per-cell dual-number blocks)
must agree with central finite differences of R along random directions, both
of this library's R and of the CPU oracle's (the reference's arithmetic
restated), and so must v^T dR/d(P0, T0) and v^T dR/d(node X, Y).  States are the fixtures' final
states (developed flows: at a uniform cold start the turbulence production is
exactly 0 and its square root has no derivative).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import PKG, REPO

pytestmark = pytest.mark.gpu

CASES = ["t2d_chan", "t2d_ramp", "taxi_duct", "t7_plate", "case1_axi"]
H = 1e-7   # relative step of the central differences


def directional(res, c, w, v, h=H):
    return float(np.sum(v * (res(c + h * w) - res(c - h * w)))) / (2.0 * h)


def check_fd(fd, gq, w, tol=2e-6):
    terms = gq * w
    ad = float(np.sum(terms))
    scale = float(np.sum(np.abs(terms)))
    assert abs(fd - ad) <= tol * scale, f"adjoint {ad:.9e} vs differences {fd:.9e} (scale {scale:.3e})"


@pytest.mark.parametrize("case", CASES)
def test_residual_equals_routine_path(golden, case):
    import tdt
    g = golden(case)
    c = g["final"]
    with tdt.Solver(g.mesh, g.flow) as s:
        rinit = s.cold_start()
        s.set_state(c, rinit)
        s.primitives()
        s.apply_boundary()
        s.source_terms()
        s.residual()
        want = np.stack([s.get_array(f"R{m}") for m in range(1, 6)])
        got = s.residual_of(c)
        assert np.array_equal(got, want)
        # the context's own state is untouched
        assert np.array_equal(s.get_state(), c)


@pytest.mark.parametrize("case", CASES)
def test_adjoint_matches_finite_differences(golden, case):
    import tdt
    g = golden(case)
    c = g["final"]
    rng = np.random.default_rng(11)
    with tdt.Solver(g.mesh, g.flow) as s:
        s.cold_start()
        for trial in range(3):
            v = rng.standard_normal(c.shape)
            w = c * rng.standard_normal(c.shape)
            gq, _ = s.residual_vjp(c, v)
            check_fd(directional(s.residual_of, c, w, v), gq, w)


@pytest.mark.parametrize("case", ["t2d_chan", "taxi_duct", "t7_plate"])
def test_adjoint_matches_reference_differences(golden, case):
    """Central differences of the oracle's R (the reference's arithmetic on the
    CPU) against this library's adjoint."""
    import tdt
    from oracle import Oracle
    g = golden(case)
    c = g["final"]
    o = Oracle(g.mesh, g.flow)
    o.cold_start()
    rinit = o.scalar("RINIT")

    def res_ref(q):
        o.set_state(q, rinit)
        o.primitives()
        o.apply_boundary()
        o.source_terms()
        o.residual()
        return np.stack([o.array(f"R{m}")[:g.mesh.neles] for m in range(1, 6)])

    rng = np.random.default_rng(5)
    v = rng.standard_normal(c.shape)
    w = c * rng.standard_normal(c.shape)
    with tdt.Solver(g.mesh, g.flow) as s:
        s.cold_start()
        gq, _ = s.residual_vjp(c, v)
        # the same R up to the device math library's ulps in pow / tanh
        got, want = s.residual_of(c), res_ref(c)
        for m in range(5):
            err = np.max(np.abs(got[m] - want[m])) / max(np.max(np.abs(want[m])), 1e-300)
            assert err <= 1e-12, f"R{m + 1}: {err:.2e}"
    check_fd(directional(res_ref, c, w, v), gq, w)


@pytest.mark.parametrize("case", ["t2d_chan", "taxi_duct", "t7_plate"])
def test_parameter_adjoint_matches_finite_differences(golden, case):
    """v^T dR/dP0 and v^T dR/dT0 (through the type-2 inlet's S0 and T0 and the
    type-7 inlet's TINF and profile speed); AMACHINIT does not enter R."""
    import dataclasses
    import tdt
    g = golden(case)
    c = g["final"]
    rng = np.random.default_rng(3)
    v = rng.standard_normal(c.shape)
    with tdt.Solver(g.mesh, g.flow) as s:
        s.cold_start()
        _, gp = s.residual_vjp(c, v)
        assert gp[2] == 0.0
        for k, name in ((0, "p0"), (1, "t0")):
            base = getattr(g.flow, name)
            h = 1e-6 * base
            r = []
            for sign in (1.0, -1.0):
                s.set_flow(dataclasses.replace(g.flow, **{name: base + sign * h}))
                r.append(s.residual_of(c))
            s.set_flow(g.flow)
            fd = float(np.sum(v * (r[0] - r[1]))) / (2.0 * h)
            assert abs(fd - gp[k]) <= 1e-5 * abs(gp[k]) + 1e-12, (name, fd, gp[k])


def node_scale(mesh):
    """Per node, its shortest adjacent cell edge (mm)."""
    x, y = mesh.x_mm, mesh.y_mm
    h = np.full(mesh.nods, np.inf)
    for f in range(4):
        a, b = mesh.nod[f] - 1, mesh.nod[(f + 1) % 4] - 1
        d = np.hypot(x[a] - x[b], y[a] - y[b])
        np.minimum.at(h, a, d)
        np.minimum.at(h, b, d)
    return h


@pytest.mark.parametrize("case", ["t2d_chan", "taxi_duct", "t7_plate", "case1_axi"])
def test_node_adjoint_matches_finite_differences(golden, case):
    """v^T dR/d(X, Y) through the geometry (surface vectors, volumes, cell and
    ghost centroids) against central differences of R with moved nodes."""
    import tdt
    g = golden(case)
    c = g["final"]
    rng = np.random.default_rng(7)
    x0, y0 = g.mesh.x_mm.astype(float), g.mesh.y_mm.astype(float)
    sc = node_scale(g.mesh)
    with tdt.Solver(g.mesh, g.flow) as s:
        s.cold_start()
        for trial in range(2):
            v = rng.standard_normal(c.shape)
            wx = sc * rng.standard_normal(g.mesh.nods)
            wy = sc * rng.standard_normal(g.mesh.nods)
            gq, gp, gxy = s.residual_vjp(c, v, nodes=True)
            h = 1e-7
            r = []
            for sign in (1.0, -1.0):
                s.set_nodes(x0 + sign * h * wx, y0 + sign * h * wy)
                r.append(s.residual_of(c))
            s.set_nodes(x0, y0)
            fd = float(np.sum(v * (r[0] - r[1]))) / (2.0 * h)
            check_fd(fd, gxy, np.stack([wx, wy]))
        # moving the nodes back restores the geometry bit for bit
        assert np.array_equal(s.get_array("VOL"), tdt.Solver(g.mesh, g.flow).get_array("VOL"))


TORCH_CHECK = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import tdt
from tdt import gridgen
from tdt.autograd import residual
import torch

mesh = gridgen.channel(nx=16, ny=8, bc={"left": 2, "right": 3, "bottom": 5, "top": 4})
flow = gridgen.flow_channel()
with tdt.Solver(mesh, flow) as s:
    s.cold_start()
    s.iterate(40)
    c = s.get_state()
    rng = np.random.default_rng(1)
    v = rng.standard_normal(c.shape)
    q = torch.tensor(c, requires_grad=True)
    p = torch.tensor([flow.p0, flow.t0, flow.amachinit], dtype=torch.float64, requires_grad=True)
    xy = torch.tensor(np.stack([mesh.x_mm, mesh.y_mm]), requires_grad=True)
    out = residual(q, p, s, xy)
    assert np.array_equal(out.detach().numpy(), s.residual_of(c))
    (out * torch.from_numpy(v)).sum().backward()
    gq, gp, gxy = s.residual_vjp(c, v, nodes=True)
    assert np.array_equal(q.grad.numpy(), gq), "grad q"
    assert np.array_equal(p.grad.numpy(), gp), "grad p"
    assert np.array_equal(xy.grad.numpy(), gxy), "grad xy"
    # a second use with other parameters rebuilds the flow constants
    p2 = torch.tensor([1.1 * flow.p0, flow.t0, flow.amachinit], dtype=torch.float64)
    r2 = residual(torch.tensor(c), p2, s).numpy()
    assert s.flow.p0 == 1.1 * flow.p0 and not np.array_equal(r2, out.detach().numpy())
print("torch autograd: OK", flush=True)
"""


def test_torch_autograd_function(tmp_path):
    """The torch.autograd.Function over the library: forward = R, backward = the
    adjoint.  In a process of its own (torch brings its own copy of the HIP
    runtime; the binding loads the library first)."""
    script = tmp_path / "torch_check.py"
    script.write_text(TORCH_CHECK)
    r = subprocess.run([sys.executable, str(script), PKG], capture_output=True, text=True,
                       timeout=240, cwd=REPO)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "torch autograd: OK" in r.stdout
