"""The adjoint of whole iterations on the GPU (tdt_iteration_vjp,
tdt_adjoint_forward / tdt_adjoint_reverse).

The operators it differentiates (DT, DISSIPATION's D, COMPUTE_VISCOUS's VIS)
are evaluated with the routine kernels' arithmetic, so their values must equal
the routine entry points' bit for bit.  The adjoint of one iteration and of N
iterations must agree with central finite differences of this library's
iteration (tdt_iterate, whose FP64 states the oracle and the reference pin) along
random directions in the state, the flow parameters and the node coordinates,
and the N-iteration adjoint also with central differences of the oracle (the
reference's arithmetic restated on the CPU).  States are the fixtures' final
states (developed flows; at a uniform cold start the turbulence production is
exactly 0 and its square root has no derivative).
"""
import dataclasses
import subprocess
import sys

import numpy as np
import pytest

from conftest import PKG, REPO

pytestmark = pytest.mark.gpu

CASES = ["t2d_chan", "case1", "taxi_duct", "t7_plate", "case1_axi"]   # LOCAL = 1


def check(fd, ad, scale, tol, what):
    assert abs(fd - ad) <= tol * scale, f"{what}: adjoint {ad:.9e} vs differences {fd:.9e} " \
                                        f"(scale {scale:.3e})"


def one_iteration(s, c, rinit):
    s.set_state(c, rinit)
    s.iterate(1, history=False)
    return s.get_state()


@pytest.mark.parametrize("case", CASES)
def test_operators_equal_routine_arrays(golden, case):
    import tdt
    g = golden(case)
    c = g["final"]
    with tdt.Solver(g.mesh, g.flow) as s:
        rinit = s.cold_start()
        s.set_state(c, rinit)
        s.timestep()
        s.dissipation()
        s.primitives()
        s.apply_boundary()
        s.viscous()
        got = s.operators_of(c)
        for name, v in got.items():
            assert np.array_equal(v, s.get_array(name)[:g.mesh.neles]), name


def operators(s, c, rinit):
    """(E, DT, VOL) [7, neles] with E = D + VIS (no VIS in the mass row)."""
    s.set_state(c, rinit)
    o = s.operators_of(c)
    return np.stack([o["D241"], o["D242"] + o["VISU"], o["D243"] + o["VISV"],
                     o["D244"] + o["VIST"], o["D245"] + o["VISR"], o["DT"],
                     s.get_array("VOL")[:c.shape[1]]])


def weighted_difference(w, plus, minus, h):
    """sum of w * (plus - minus) / 2h, differenced entry by entry before the sum
    (a sum of large terms would bury a small change in its rounding)."""
    return float(np.sum(w * (plus - minus))) / (2 * h)


def untied_nodes(mesh, rng, frac=1e-3):
    """Node coordinates moved by frac of their shortest edge: on a uniform grid
    DL = min(edge lengths) has tied edges, where it has no derivative."""
    sc = node_scale(mesh)
    return (mesh.x_mm + frac * sc * rng.uniform(-1, 1, mesh.nods),
            mesh.y_mm + frac * sc * rng.uniform(-1, 1, mesh.nods))


@pytest.mark.parametrize("case", CASES)
def test_operator_adjoint_matches_finite_differences(golden, case):
    """w^T d(D + VIS, DT)/dQ, w^T d(D + VIS)/d(P0, T0) and w^T d(D + VIS, DT,
    VOL)/d(X, Y) of tdt_operators_vjp against central differences of
    tdt_operators_eval (weights scaled per row, nodes moved off ties)."""
    import tdt
    g = golden(case)
    c = g["final"]
    rng = np.random.default_rng(17)
    with tdt.Solver(g.mesh, g.flow) as s:
        x0, y0 = untied_nodes(g.mesh, rng)
        s.set_nodes(x0, y0)
        rinit = s.cold_start()
        e0 = operators(s, c, rinit)
        w = rng.standard_normal(e0.shape) / np.maximum(np.abs(e0).max(axis=1, keepdims=True),
                                                       1e-300)
        gq, gp, gxy = s.operators_vjp(c, rinit, w, nodes=True)
        dq = c * rng.standard_normal(c.shape)
        # small: the sensors' |p_j - p_i| and max() have kinks a larger step
        # can cross where neighbours are nearly equal (across a boundary layer)
        h = 1e-9
        fd = weighted_difference(w, operators(s, c + h * dq, rinit),
                                 operators(s, c - h * dq, rinit), h)
        terms = gq * dq
        check(fd, float(np.sum(terms)), float(np.sum(np.abs(terms))), 1e-6, "state")
        for k, name in ((0, "p0"), (1, "t0")):
            base = getattr(g.flow, name)
            hp = 1e-6 * base
            e = []
            for sign in (1.0, -1.0):
                s.set_flow(dataclasses.replace(g.flow, **{name: base + sign * hp}))
                e.append(operators(s, c, rinit))
            s.set_flow(g.flow)
            check(weighted_difference(w, e[0], e[1], hp), gp[k], 2 * abs(gp[k]) + 1e-300, 1e-5,
                  name)
        sc = node_scale(g.mesh)
        wx = sc * rng.standard_normal(g.mesh.nods)
        wy = sc * rng.standard_normal(g.mesh.nods)
        h = 1e-7
        e = []
        for sign in (1.0, -1.0):
            s.set_nodes(x0 + sign * h * wx, y0 + sign * h * wy)
            e.append(operators(s, c, rinit))
        s.set_nodes(x0, y0)
        fd = weighted_difference(w, e[0], e[1], h)
        terms = gxy * np.stack([wx, wy])
        # 1e-5: on case1 the differences of D + VIS in the nodes settle only to a
        # few 1e-6 of the scale (steps 1e-6..1e-8 give 3e-5, 4e-6, 3e-6 apart,
        # not converging: rounding of terms far larger than their sum), while DT,
        # VOL and every row on the other grids agree to 1e-8 or better
        check(fd, float(np.sum(terms)), float(np.sum(np.abs(terms))), 1e-5, "nodes")


@pytest.mark.parametrize("case", ["t2d_chan", "t7_plate", "taxi_duct"])
def test_operator_rinit_adjoint_with_the_nu_floor_active(golden, case):
    """w^T d(D + VIS)/dRINIT with RTRED = 2 (RTOL above every nu~: the sensor's
    and the node averages' RTOL bounds decide, and RTOL's effect stands far
    above rounding; at the flow.in RTRED it moves E by ~1e-11 relative)."""
    import tdt
    g = golden(case)
    c = g["final"]
    rng = np.random.default_rng(19)
    with tdt.Solver(g.mesh, dataclasses.replace(g.flow, rtred=2.0)) as s:
        rinit = s.cold_start()
        e0 = operators(s, c, rinit)
        w = rng.standard_normal(e0.shape) / np.maximum(np.abs(e0).max(axis=1, keepdims=True),
                                                       1e-300)
        _, gp = s.operators_vjp(c, rinit, w)
        hr = 1e-5 * rinit
        fd = weighted_difference(w, operators(s, c, rinit + hr), operators(s, c, rinit - hr), hr)
        check(fd, gp[2], 2 * abs(gp[2]), 1e-6, "rinit")


@pytest.mark.parametrize("case", CASES)
def test_iteration_adjoint_matches_finite_differences(golden, case):
    """mu^T dQ^{n+1}/dQ^n along random state directions, and the parameter
    derivatives (P0, T0 through the flow constants, RINIT through RTOL)."""
    import tdt
    g = golden(case)
    c = g["final"]
    rng = np.random.default_rng(21)
    with tdt.Solver(g.mesh, g.flow) as s:
        rinit = s.cold_start()
        for trial in range(2):
            mu = rng.standard_normal(c.shape) / np.abs(c).max(axis=1, keepdims=True)
            w = c * rng.standard_normal(c.shape)
            lam, gp = s.iteration_vjp(c, rinit, mu)
            h = 1e-9   # below the sensors' kinks (see the operator test)
            fd = float(np.sum(mu * (one_iteration(s, c + h * w, rinit)
                                    - one_iteration(s, c - h * w, rinit)))) / (2 * h)
            terms = lam * w
            check(fd, float(np.sum(terms)), float(np.sum(np.abs(terms))), 1e-6, "state")
        # the context's state is left as it was
        s.set_state(c, rinit)
        before = s.get_state()
        s.iteration_vjp(c, rinit, mu)
        assert np.array_equal(s.get_state(), before)
        for k, name in ((0, "p0"), (1, "t0")):
            base = getattr(g.flow, name)
            h = 1e-6 * base
            q = []
            for sign in (1.0, -1.0):
                s.set_flow(dataclasses.replace(g.flow, **{name: base + sign * h}))
                q.append(one_iteration(s, c, rinit))
            s.set_flow(g.flow)
            fd = float(np.sum(mu * (q[0] - q[1]))) / (2 * h)
            check(fd, gp[k], abs(gp[k]) + abs(fd) + 1e-300, 1e-5, name)
        # RINIT enters through RTOL = RINIT * RTRED (~1e-10 here), whose effect on
        # one iteration lies at the state's rounding level: it is checked at the
        # operator level above and with a raised floor below


@pytest.mark.parametrize("case", ["t2d_chan", "t7_plate", "taxi_duct"])
def test_iteration_adjoint_with_the_nu_floor_active(golden, case):
    """RTRED = 2 puts RTOL = RINIT * RTRED above every nu~ of the state, so each
    max() RTOL enters is decided for RTOL: the floor C5 = C1 max(RTOL, C5/C1)
    clamps every cell at every stage, the dissipation's nu~ sensor takes its
    8 C1 RTOL bound and the node averages their RTOL bound.  RTOL's effect is
    then smooth and far above rounding: the RINIT derivative and the state
    adjoint against central differences (weights scaled per equation)."""
    import tdt
    g = golden(case)
    c = g["final"]
    flow = dataclasses.replace(g.flow, rtred=2.0)
    rng = np.random.default_rng(31)
    with tdt.Solver(g.mesh, flow) as s:
        rinit = s.cold_start()
        assert np.all(c[4] / c[0] < rinit * flow.rtred)
        mu = rng.standard_normal(c.shape) / np.abs(c).max(axis=1, keepdims=True)
        lam, gp = s.iteration_vjp(c, rinit, mu)
        h = 1e-4 * rinit
        fd = float(np.sum(mu * (one_iteration(s, c, rinit + h)
                                - one_iteration(s, c, rinit - h)))) / (2 * h)
        check(fd, gp[2], abs(gp[2]) + abs(fd), 1e-6, "rinit")
        w = c * rng.standard_normal(c.shape)
        h = 1e-7
        fd = float(np.sum(mu * (one_iteration(s, c + h * w, rinit)
                                - one_iteration(s, c - h * w, rinit)))) / (2 * h)
        terms = lam * w
        check(fd, float(np.sum(terms)), float(np.sum(np.abs(terms))), 1e-6, "state")


def node_scale(mesh):
    x, y = mesh.x_mm, mesh.y_mm
    h = np.full(mesh.nods, np.inf)
    for f in range(4):
        a, b = mesh.nod[f] - 1, mesh.nod[(f + 1) % 4] - 1
        d = np.hypot(x[a] - x[b], y[a] - y[b])
        np.minimum.at(h, a, d)
        np.minimum.at(h, b, d)
    return h


@pytest.mark.parametrize("case", ["t2d_chan", "taxi_duct", "t7_plate", "case1_axi"])
def test_iteration_node_adjoint_matches_finite_differences(golden, case):
    """mu^T dQ^{n+1}/d(X, Y) through every geometric quantity one iteration
    reads (surface vectors, volumes, centroids, ghost centroids, DL)."""
    import tdt
    g = golden(case)
    c = g["final"]
    rng = np.random.default_rng(9)
    sc = node_scale(g.mesh)
    with tdt.Solver(g.mesh, g.flow) as s:
        x0, y0 = untied_nodes(g.mesh, rng)
        s.set_nodes(x0, y0)
        rinit = s.cold_start()
        mu = rng.standard_normal(c.shape) / np.abs(c).max(axis=1, keepdims=True)
        _, _, gxy = s.iteration_vjp(c, rinit, mu, nodes=True)
        wx = sc * rng.standard_normal(g.mesh.nods)
        wy = sc * rng.standard_normal(g.mesh.nods)
        h = 1e-7
        q = []
        for sign in (1.0, -1.0):
            s.set_nodes(x0 + sign * h * wx, y0 + sign * h * wy)
            q.append(one_iteration(s, c, rinit))
        s.set_nodes(x0, y0)
        fd = float(np.sum(mu * (q[0] - q[1]))) / (2 * h)
        terms = gxy * np.stack([wx, wy])
        check(fd, float(np.sum(terms)), float(np.sum(np.abs(terms))), 1e-6, "nodes")


def unrolled_case(golden, name):
    g = golden(name)
    return g.mesh, g.flow


@pytest.mark.parametrize("case", ["t2d_chan", "taxi_duct", "t7_plate"])
def test_unrolled_adjoint_matches_finite_differences(golden, case):
    """J = v^T Q^N after N iterations from the cold start: dJ/d(P0, T0,
    AMACHINIT) (through the flow constants, the cold-start state and RINIT)
    and dJ/dQ^0 along a random direction, against central differences of
    tdt_iterate runs."""
    import tdt
    mesh, flow = unrolled_case(golden, case)
    N = 30
    rng = np.random.default_rng(4)
    with tdt.Solver(mesh, flow) as s:
        qn = s.adjoint_forward(N)
        # the forward run is tdt_iterate's
        rinit = s.cold_start()
        s.iterate(N, history=False)
        assert np.array_equal(qn, s.get_state())
        v = rng.standard_normal(qn.shape) / np.abs(qn).max(axis=1, keepdims=True)
        lam0, gp = s.adjoint_reverse(v)
        for k, name in ((0, "p0"), (1, "t0"), (2, "amachinit")):
            base = getattr(flow, name)
            h = 1e-6 * base
            j = []
            for sign in (1.0, -1.0):
                s.set_flow(dataclasses.replace(flow, **{name: base + sign * h}))
                s.cold_start()
                s.iterate(N, history=False)
                j.append(float(np.sum(v * s.get_state())))
            s.set_flow(flow)
            fd = (j[0] - j[1]) / (2 * h)
            check(fd, gp[k], abs(gp[k]) + abs(fd), 1e-4, name)
        # from a given start state: dJ/dQ^0 along w
        q0 = golden(case)["final"]
        s.adjoint_forward(N, q0, rinit)
        lam0, gp = s.adjoint_reverse(v)
        w = q0 * rng.standard_normal(q0.shape)
        h = 1e-7
        j = []
        for sign in (1.0, -1.0):
            s.set_state(q0 + sign * h * w, rinit)
            s.iterate(N, history=False)
            j.append(float(np.sum(v * s.get_state())))
        fd = (j[0] - j[1]) / (2 * h)
        terms = lam0 * w
        check(fd, float(np.sum(terms)), float(np.sum(np.abs(terms))), 1e-5, "Q0")
        assert gp[2] == 0.0


@pytest.mark.parametrize("case", ["t2d_chan", "taxi_duct"])
def test_unrolled_adjoint_matches_reference_differences(golden, case):
    """The same parameter gradients against central differences of the oracle
    (the reference's arithmetic on the CPU) from its own cold start."""
    import tdt
    from oracle import Oracle
    mesh, flow = unrolled_case(golden, case)
    N = 20
    rng = np.random.default_rng(8)
    with tdt.Solver(mesh, flow) as s:
        qn = s.adjoint_forward(N)
        v = rng.standard_normal(qn.shape) / np.abs(qn).max(axis=1, keepdims=True)
        _, gp = s.adjoint_reverse(v)
    for k, name in ((0, "p0"), (1, "t0"), (2, "amachinit")):
        base = getattr(flow, name)
        h = 1e-6 * base
        j = []
        for sign in (1.0, -1.0):
            o = Oracle(mesh, dataclasses.replace(flow, **{name: base + sign * h}))
            o.cold_start()
            o.iterate(N)
            j.append(float(np.sum(v * o.state())))
        fd = (j[0] - j[1]) / (2 * h)
        check(fd, gp[k], abs(gp[k]) + abs(fd), 1e-4, name)


def test_unrolled_node_adjoint(golden):
    """dJ/d(X, Y) after N iterations from the cold start."""
    import tdt
    mesh, flow = unrolled_case(golden, "taxi_duct")
    N = 15
    rng = np.random.default_rng(12)
    sc = node_scale(mesh)
    with tdt.Solver(mesh, flow) as s:
        x0, y0 = untied_nodes(mesh, rng)
        s.set_nodes(x0, y0)
        qn = s.adjoint_forward(N)
        v = rng.standard_normal(qn.shape) / np.abs(qn).max(axis=1, keepdims=True)
        _, _, gxy = s.adjoint_reverse(v, nodes=True)
        wx = sc * rng.standard_normal(mesh.nods)
        wy = sc * rng.standard_normal(mesh.nods)
        h = 1e-7
        j = []
        for sign in (1.0, -1.0):
            s.set_nodes(x0 + sign * h * wx, y0 + sign * h * wy)
            s.cold_start()
            s.iterate(N, history=False)
            j.append(float(np.sum(v * s.get_state())))
        s.set_nodes(x0, y0)
        fd = (j[0] - j[1]) / (2 * h)
        terms = gxy * np.stack([wx, wy])
        check(fd, float(np.sum(terms)), float(np.sum(np.abs(terms))), 1e-5, "nodes")


def test_unrolled_segments_match_one_segment(golden):
    """The reverse sweep's checkpoint segments (every ceil(sqrt(N)) iterations)
    give the adjoint a sweep over one segment gives: N = 9 (3 segments of 3)
    against 9 single-iteration reverses chained by hand."""
    import tdt
    mesh, flow = unrolled_case(golden, "t2d_chan")
    N = 9
    rng = np.random.default_rng(2)
    with tdt.Solver(mesh, flow) as s:
        rinit = s.cold_start()
        states = [s.get_state()]
        for _ in range(N):
            s.iterate(1, history=False)
            states.append(s.get_state())
        s.adjoint_forward(N, states[0], rinit)
        v = rng.standard_normal(states[-1].shape)
        lam0, gp = s.adjoint_reverse(v)
        lam, gsum = v, np.zeros(3)
        for it in range(N - 1, -1, -1):
            lam, g1 = s.iteration_vjp(states[it], rinit, lam)
            gsum += g1
        assert np.allclose(lam0, lam, rtol=1e-12, atol=1e-12 * np.abs(lam).max())
        assert np.allclose(gp[:2], gsum[:2], rtol=1e-11)
        assert np.isclose(gp[3], gsum[2], rtol=1e-11)


def test_global_time_step_is_refused(golden):
    """LOCAL = 0 (one DTMIN for all cells, an argmin) is not differentiated:
    the entry points say so instead of returning a wrong adjoint."""
    import tdt
    from tdt import lib
    g = golden("t2d_ramp")
    assert g.flow.local == 0
    with tdt.Solver(g.mesh, g.flow) as s:
        rinit = s.cold_start()
        c = g["final"]
        lam = np.empty_like(c)
        rc = s.L.tdt_iteration_vjp(s.h, lib._d(np.ascontiguousarray(c)), rinit,
                                   lib._d(np.ascontiguousarray(c)), lib._d(lam), None, None)
        assert rc == lib.TDT_ERR_UNSUPPORTED


TORCH_CHECK = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import tdt
from tdt import gridgen
from tdt.autograd import unrolled
import torch

mesh = gridgen.channel(nx=16, ny=8, bc={"left": 2, "right": 3, "bottom": 5, "top": 4})
flow = gridgen.flow_channel()
with tdt.Solver(mesh, flow) as s:
    N = 12
    p = torch.tensor([flow.p0, flow.t0, flow.amachinit], dtype=torch.float64, requires_grad=True)
    xy = torch.tensor(np.stack([mesh.x_mm, mesh.y_mm]), requires_grad=True)
    qn = unrolled(p, s, N, xy)
    s.cold_start()
    s.iterate(N, history=False)
    assert np.array_equal(qn.detach().numpy(), s.get_state()), "forward"
    rng = np.random.default_rng(1)
    v = rng.standard_normal(qn.shape)
    (qn * torch.from_numpy(v)).sum().backward()
    s.adjoint_forward(N)
    _, gp, gxy = s.adjoint_reverse(v, nodes=True)
    assert np.array_equal(p.grad.numpy(), gp[:3]), "grad p"
    assert np.array_equal(xy.grad.numpy(), gxy), "grad xy"
    # a backward after other parameters were set runs its own forward again
    q2 = unrolled(p, s, N, xy)
    s.set_flow(flow.__class__(**{**flow.__dict__, "p0": 1.1 * flow.p0}))
    p.grad = None
    (q2 * torch.from_numpy(v)).sum().backward()
    assert np.array_equal(p.grad.numpy(), gp[:3]), "grad p after set_flow"
    # the C call refuses a reverse once the parameters changed under it
    s.adjoint_forward(N)
    s.set_flow(flow)
    try:
        s.adjoint_reverse(v)
        raise SystemExit("reverse after set_flow was not refused")
    except tdt.TdtError as exc:
        assert exc.code == tdt.lib.TDT_ERR_UNSUPPORTED
print("torch unrolled: OK", flush=True)
"""


def test_torch_unrolled_function(tmp_path):
    """tdt.autograd.unrolled: Q^N of N iterations from the cold start of
    (P0, T0, AMACHINIT) with the node coordinates; backward = the reverse sweep.
    In a process of its own (the binding loads the library before torch)."""
    script = tmp_path / "torch_unrolled.py"
    script.write_text(TORCH_CHECK)
    r = subprocess.run([sys.executable, str(script), PKG], capture_output=True, text=True,
                       timeout=240, cwd=REPO)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "torch unrolled: OK" in r.stdout


def test_duct_design_gradients():
    """The 193x65 duct with 8 Hicks-Henne bumps: J = axial wall-pressure force +
    outlet mass flow after N iterations from the cold start; dJ/dP0,
    dJ/dAMACHINIT and dJ/d(bump amplitude) of two bumps from one reverse sweep
    against central differences of tdt_iterate runs."""
    import dataclasses
    import tdt
    from tdt import gridgen, objective
    mesh, flow = gridgen.WORKLOADS["duct_193x65"][0](), gridgen.WORKLOADS["duct_193x65"][1]()
    dnode = objective.bump_node_derivatives(8, nx=193, ny=65, jitter=0.01 * 193 / 385)
    N = 30

    def run_j(s, m, f):
        s.cold_start()
        s.iterate(N, history=False)
        return objective.duct_objective(m, f, s.get_state())[0]

    with tdt.Solver(mesh, flow) as s:
        qn = s.adjoint_forward(N)
        _, djdq, djdxy = objective.duct_objective(mesh, flow, qn)
        _, gp, gxy = s.adjoint_reverse(djdq, nodes=True)
        dxy = gxy + djdxy
        for k, name in ((0, "p0"), (2, "amachinit")):
            base = getattr(flow, name)
            h = 1e-6 * base
            j = []
            for sign in (1.0, -1.0):
                s.set_flow(dataclasses.replace(flow, **{name: base + sign * h}))
                j.append(run_j(s, mesh, flow))
            s.set_flow(flow)
            check((j[0] - j[1]) / (2 * h), gp[k], abs(gp[k]), 1e-5, name)
        for k in (1, 6):
            ad = float(np.sum(dxy * dnode[k]))
            h = 1e-3
            j = []
            for sign in (1.0, -1.0):
                x, y = mesh.x_mm + sign * h * dnode[k, 0], mesh.y_mm + sign * h * dnode[k, 1]
                s.set_nodes(x, y)
                j.append(run_j(s, dataclasses.replace(mesh, x_mm=x, y_mm=y), flow))
            s.set_nodes(mesh.x_mm, mesh.y_mm)
            check((j[0] - j[1]) / (2 * h), ad, abs(ad), 1e-5, f"bump{k + 1}")


def uniform_channel(p0=1.2e5, t0=298.0, mach=0.3):
    """A channel whose cold start is a steady state of the iteration: total
    conditions at the inlet (type 2), the uniform flow's static pressure at the
    outlet (type 3 with PAMB), symmetry above and below (type 4)."""
    from tdt import gridgen
    g = 1.4
    mesh = gridgen.channel(nx=16, ny=8, length=100.0, height=20.0, h1=2.0, jitter=0.02,
                           bc={"left": 2, "right": 3, "bottom": 4, "top": 4})
    pamb = p0 / (1.0 + 0.5 * (g - 1.0) * mach * mach) ** (g / (g - 1.0))
    return mesh, gridgen.flow_channel(p0=p0, t0=t0, amachinit=mach, pamb=pamb)


def steady_outlet_mass_flow(mesh, flow, p0, t0):
    """Outlet mass flow of the steady state for (P0, T0) at the flow's PAMB:
    uniform flow at the Mach number whose static pressure from P0 is PAMB."""
    from tdt import objective
    g, rr = flow.gama, 8314.0 / flow.amol
    m2 = 2.0 / (g - 1.0) * ((p0 / flow.pamb) ** ((g - 1.0) / g) - 1.0)
    t = t0 / (1.0 + 0.5 * (g - 1.0) * m2)
    rho = flow.pamb / (rr * t)
    u = np.sqrt(m2 * g * rr * t)
    out = objective.outlet_faces(mesh)
    sx, _, _ = objective._scx(mesh, 0, out[:, 1], out[:, 2])
    return float(rho * u * np.sum(sx))


def test_fixed_point_graph_spans_equal_eager_iterations(monkeypatch):
    """The fixed-point loop runs the iterations between its convergence checks as a
    captured graph, replayed many times with eager iterations and host reads in
    between; launched one by one (TDT_FIXED_POINT_EAGER) the same kernels give the
    same adjoint, parameter and node gradients, iteration count and error estimate,
    bit for bit.  Odd max_iter: the run ends in a partial span, eagerly."""
    import tdt
    from tdt import objective
    mesh, flow = uniform_channel()
    out = {}
    for mode in ("graph", "eager"):
        if mode == "eager":
            monkeypatch.setenv("TDT_FIXED_POINT_EAGER", "1")
        with tdt.Solver(mesh, flow) as s:
            rinit = s.cold_start()
            c = s.get_state()
            o = objective.outlet_faces(mesh)
            sx, _, _ = objective._scx(mesh, 0, o[:, 1], o[:, 2])
            dj = np.zeros_like(c)
            np.add.at(dj[1], o[:, 0], sx)
            out[mode] = s.adjoint_fixed_point(c, rinit, dj, max_iter=1237, tol=0.0, nodes=True)
    (lg, gg, xg, ig, cg), (le, ge, xe, ie, ce) = out["graph"], out["eager"]
    assert ig == ie and cg == ce
    assert np.array_equal(lg, le) and np.array_equal(gg, ge) and np.array_equal(xg, xe)


def test_fixed_point_adjoint_at_a_steady_state():
    """The adjoint iterated to its fixed point, lam = Phi_Q^T lam + dJ/dQ, at a
    steady state of the iteration (the uniform channel): J = outlet mass flow;
    dJ/dP0 = lam^T dPhi/dP0 must equal the derivative of the steady state's mass
    flow in closed form (the outlet pressure pins the Mach number), dJ/dT0 too,
    and RINIT (which only moves inactive floors) contributes nothing."""
    import tdt
    from tdt import objective
    mesh, flow = uniform_channel()
    with tdt.Solver(mesh, flow) as s:
        rinit = s.cold_start()
        c = s.get_state()
        s.iterate(50, history=False)
        scale = np.abs(c).max(axis=1)
        scale[2] = scale[1]                          # rho v is 0: measure it against rho u
        drift = np.max(np.abs(s.get_state() - c) / scale[:, None])
        assert drift < 1e-12, drift                  # a fixed point of the iteration
        out = objective.outlet_faces(mesh)
        sx, _, _ = objective._scx(mesh, 0, out[:, 1], out[:, 2])
        dj = np.zeros_like(c)
        np.add.at(dj[1], out[:, 0], sx)              # J = sum (rho u) SC_x at the outlet
        assert abs(float(np.sum(dj * c)) - steady_outlet_mass_flow(mesh, flow, flow.p0,
                                                                   flow.t0)) < 1e-9
        lam, gp, iters, change = s.adjoint_fixed_point(c, rinit, dj, max_iter=60000, tol=1e-12)
        assert change <= 1e-12, (iters, change)
        for k, name in ((0, "p0"), (1, "t0")):
            base = getattr(flow, name)
            h = 1e-6 * base
            args = [dict(p0=flow.p0, t0=flow.t0) for _ in range(2)]
            args[0][name] = base + h
            args[1][name] = base - h
            fd = (steady_outlet_mass_flow(mesh, flow, **args[0])
                  - steady_outlet_mass_flow(mesh, flow, **args[1])) / (2 * h)
            check(fd, gp[k], abs(fd), 1e-6, name)
        assert abs(gp[2]) * rinit < 1e-12 * abs(gp[0]) * flow.p0


CONVERGED = "channel_48x16_converged"


@pytest.mark.skipif(not __import__("os").path.exists(__import__("os").path.join(
    REPO, "tests", "golden", CONVERGED + ".npz")), reason=CONVERGED + " fixture not generated")
def test_fixed_point_adjoint_at_a_converged_nonuniform_flow(golden):
    """The fixed-point adjoint at a developed, non-uniform steady state: the
    48 x 16 channel (total-condition inlet, outlet at a fixed static pressure,
    no-slip wall, symmetry) converged from the cold start to an L2 residual drop
    below 1e-8 (a boundary layer grows along the wall; the state is nowhere
    uniform).  J = outlet mass flow.  The steady state's dJ/dP0 and dJ/dT0 from
    lam^T dPhi/dp (the cold start's RINIT = f(P0, T0) chained in) against
    central differences (relative step 1e-6) of converged runs of this library
    (the five runs as one batch) and of the reference solver (tests/golden: its
    converged runs at P0, T0 moved by +-1e-6), within 1e-4.  The converged
    state's J equals the reference's to 1e-9."""
    import dataclasses as dc
    import tdt
    from tdt import objective
    g = golden(CONVERGED)
    mesh, flow, n = g.mesh, g.flow, g.nrun
    out = objective.outlet_faces(mesh)
    sx, _, _ = objective._scx(mesh, 0, out[:, 1], out[:, 2])

    def J(c):
        return float(np.sum(c[1, out[:, 0]] * sx))

    # the converged state and the four moved runs as one batch (five cases in one
    # fused iteration; each case's state is that of a context of its own, bit for bit)
    moved = [(par, sg) for par in ("p0", "t0") for sg in (1.0, -1.0)]
    cases = [(mesh, flow)] + [(mesh, dc.replace(flow, **{par: getattr(flow, par) * (1.0 + sg * 1e-6)}))
                              for par, sg in moved]
    with tdt.Batch(cases) as b:
        b.cold_start()
        l2a, _, bad = b.iterate(50)
        assert bad == 0
        b.iterate(n - 51, history=False)
        l2b, _, bad = b.iterate(1)
        assert bad == 0
        drop = float(np.max(l2b[-1, 0] / l2a[:, 0].max(axis=0)))
        c = b.get_state(0)
        jm = {key: J(b.get_state(k + 1)) for k, key in enumerate(moved)}
    j = J(c)
    assert drop <= 1e-8, drop
    assert abs(j - float(g["J_base"])) <= 1e-9 * abs(float(g["J_base"])), (j, g["J_base"])
    fd = {par: (jm[(par, 1.0)] - jm[(par, -1.0)]) / (2e-6 * getattr(flow, par))
          for par in ("p0", "t0")}
    with tdt.Solver(mesh, flow) as s:
        rinit = s.cold_start()
        dj = np.zeros_like(c)
        np.add.at(dj[1], out[:, 0], sx)
        _, gp, iters, est = s.adjoint_fixed_point(c, rinit, dj, max_iter=1500000, tol=1e-8)
        assert est <= 1e-8, (iters, est)
    # RINIT = AMUINIT / RHOINIT of the cold start: proportional to T0^1.6 / P0
    adj = {"p0": gp[0] - gp[2] * rinit / flow.p0,
           "t0": gp[1] + gp[2] * 1.6 * rinit / flow.t0}
    ref = {par: (float(g[f"J_{par}+1e-06"]) - float(g[f"J_{par}-1e-06"]))
           / (2e-6 * getattr(flow, par)) for par in ("p0", "t0")}
    print(f"converged to an L2 drop of {drop:.2e}; adjoint {iters} iterations (error estimate "
          f"{est:.1e}); dJ/dp adjoint {adj}, differences {fd}, reference differences {ref}")
    for par in ("p0", "t0"):
        check(fd[par], adj[par], abs(fd[par]), 1e-4, par + " vs differences")
        check(ref[par], adj[par], abs(ref[par]), 1e-4, par + " vs reference differences")


DUCT_CONVERGED = "duct_193x65_converged"
DUCT_NRUN = 850000


@pytest.fixture(scope="module")
def converged_duct():
    """The adjoint configuration's 193 x 65 duct (8 Hicks-Henne bumps at zero
    amplitude; CFL 0.5, at which its flow settles: at the specified 0.25 it keeps
    oscillating) run from the cold start for DUCT_NRUN iterations in one batch with
    its variants at P0, T0 moved by +-1e-6 relative and bump 3 at +-1e-3 mm; then
    the fixed-point adjoint of J = axial wall-pressure force + outlet mass flow at
    the base case's steady state, with the node gradient chained through the
    generator's dX/da_3."""
    import dataclasses as dc
    import tdt
    from tdt import gridgen, objective
    duct_kw = dict(nx=193, ny=65, jitter=0.01 * 193 / 385)
    mesh, flow = gridgen.duct(bumps=[0.0] * 8, **duct_kw), gridgen.flow_duct(cfl=0.5)
    cases, tags = [(mesh, flow)], ["base"]
    for par in ("p0", "t0"):
        for sg in (1.0, -1.0):
            cases.append((mesh, dc.replace(flow, **{par: getattr(flow, par) * (1.0 + sg * 1e-6)})))
            tags.append(par + ("+" if sg > 0 else "-"))
    for sg in (1.0, -1.0):
        amps = [0.0] * 8
        amps[2] = sg * 1e-3
        cases.append((gridgen.duct(bumps=amps, **duct_kw), flow))
        tags.append("bump3" + ("+" if sg > 0 else "-"))
    with tdt.Batch(cases) as b:
        b.cold_start()
        l2a, _, bad = b.iterate(50)
        assert bad == 0
        b.iterate(DUCT_NRUN - 51, history=False)
        l2b, _, bad = b.iterate(1)
        assert bad == 0
        drop = (l2b[-1] / l2a.max(axis=0)).max(axis=1)
        states = [b.get_state(k) for k in range(len(cases))]
    J = {t: objective.duct_objective(m, f, c)[0] for t, (m, f), c in zip(tags, cases, states)}
    fd = {par: (J[par + "+"] - J[par + "-"]) / (2e-6 * getattr(flow, par)) for par in ("p0", "t0")}
    fd["bump3"] = (J["bump3+"] - J["bump3-"]) / 2e-3
    j, djdq, djdxy = objective.duct_objective(mesh, flow, states[0])
    with tdt.Solver(mesh, flow) as s:
        rinit = s.cold_start()
        _, gp, gxy, iters, est = s.adjoint_fixed_point(states[0], rinit, djdq, max_iter=2000000,
                                                       tol=1e-9, nodes=True)
    # RINIT = AMUINIT / RHOINIT of the cold start: proportional to T0^1.6 / P0
    adj = {"p0": gp[0] - gp[2] * rinit / flow.p0, "t0": gp[1] + gp[2] * 1.6 * rinit / flow.t0,
           "bump3": float(np.sum((gxy + djdxy) * objective.bump_node_derivatives(8, **duct_kw)[2]))}
    return dict(flow=flow, drop=drop, J=j, fd=fd, adj=adj, iters=iters, est=est, state=states[0])


def test_fixed_point_adjoint_on_the_converged_adjoint_duct(converged_duct):
    """The 193 x 65 duct converged to an L2 residual drop below 1e-8 (every case of
    the batch); the fixed-point adjoint's dJ/dP0, dJ/dT0 and dJ/d(bump 3) of the
    steady state against central differences of the converged runs, within 1e-4."""
    r = converged_duct
    assert np.all(r["drop"] <= 1e-8), r["drop"]
    assert r["est"] <= 1e-9, (r["iters"], r["est"])
    print(f"duct converged: L2 drop {r['drop'].max():.2e}, J = {r['J']:.12e}; adjoint "
          f"{r['iters']} iterations (error estimate {r['est']:.1e}); adjoint {r['adj']}, "
          f"differences {r['fd']}")
    for key in ("p0", "t0", "bump3"):
        check(r["fd"][key], r["adj"][key], abs(r["fd"][key]), 1e-4, key + " vs differences")


@pytest.mark.skipif(not __import__("os").path.exists(__import__("os").path.join(
    REPO, "tests", "golden", DUCT_CONVERGED + ".npz")), reason=DUCT_CONVERGED + " fixture not generated")
def test_converged_adjoint_duct_against_the_reference(converged_duct, golden):
    """The reference solver run on the same duct for the same iterations
    (tests/golden): the converged J agrees to 1e-9, and the adjoint's gradients
    agree with the reference's central differences (P0, T0 by +-1e-6 relative,
    bump 3 by +-1e-3 mm) within 1e-4."""
    r = converged_duct
    g = golden(DUCT_CONVERGED)
    assert int(g.nrun) == DUCT_NRUN
    jb = float(g["J_base"])
    assert abs(r["J"] - jb) <= 1e-9 * abs(jb), (r["J"], jb)
    h, hb = float(g["h_par"]), float(g["h_bump"])
    ref = {par: (float(g[f"J_{par}+"]) - float(g[f"J_{par}-"])) / (2 * h * getattr(r["flow"], par))
           for par in ("p0", "t0")}
    ref["bump3"] = (float(g["J_bump3+"]) - float(g["J_bump3-"])) / (2 * hb)
    print(f"reference differences {ref}, adjoint {r['adj']}")
    for key in ("p0", "t0", "bump3"):
        check(ref[key], r["adj"][key], abs(ref[key]), 1e-4, key + " vs reference differences")
