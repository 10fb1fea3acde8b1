"""Structured quadrilateral grids written in the solver's grid.in / bc.in form.

Numbering follows examples/case1 of the reference:
  node n(i, j) = i*(ny+1) + j + 1      (i along x, j along y, j fastest)
  cell e(i, j) = i*ny + j + 1
  cell nodes (i,j), (i+1,j), (i+1,j+1), (i,j+1)  -> counter-clockwise
  face k joins nodes k and k+1: 1 bottom, 2 right, 3 top, 4 left
  NCk = neighbour across face k, or the id of the ghost cell on that face
        (geometry.f90:41-42; ghost ids are NELES+1.. in the order cells and
        then faces 1..4 are visited, as in examples/case1/bc.in)

Boundary types: 2 subsonic inlet, 3 extrapolation "farfield", 4 symmetry,
5 no-slip wall / axis, 7 1/7-power boundary-layer inlet (boundary.f90:24-48).

Coordinates are rounded to the 10 decimals grid.in carries, so a grid passed
in memory and the same grid read back from its text are bit-identical.
Transcendental functions are evaluated with the math module (libm) on 1-D
profiles only; 2-D arithmetic is +,-,*,/ in numpy, which is exactly rounded,
so a seed always gives the same grid on any host.
"""
from __future__ import annotations

import math

import numpy as np

from .caseio import Flow, Mesh

DEFAULT_SEED = 20261203


def _round10(a: np.ndarray) -> np.ndarray:
    return np.array([float("%.10f" % v) for v in np.asarray(a, dtype=np.float64).ravel()]).reshape(np.shape(a))


def geometric_ratio(h1: float, total: float, n: int) -> float:
    """q such that h1*(q**n - 1)/(q - 1) == total (bisection, q >= 1)."""
    if h1 * n >= total:
        return 1.0
    lo, hi = 1.0 + 1e-15, 2.0
    while h1 * (hi ** n - 1.0) / (hi - 1.0) < total:
        hi *= 2.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if h1 * (mid ** n - 1.0) / (mid - 1.0) < total:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def stretched(h1: float, total: float, n: int) -> np.ndarray:
    """n+1 stations from 0 to total, first spacing h1, geometric growth."""
    q = geometric_ratio(h1, total, n)
    s = np.zeros(n + 1)
    h = h1
    for j in range(1, n + 1):
        s[j] = s[j - 1] + h
        h *= q
    s *= total / s[n]
    s[n] = total
    return s


def structured_mesh(xn: np.ndarray, yn: np.ndarray, bc: dict) -> Mesh:
    """Mesh from node coordinates xn, yn of shape [nx+1, ny+1] (mm).

    bc maps 'bottom', 'right', 'top', 'left' to boundary types.
    """
    nx, ny = xn.shape[0] - 1, xn.shape[1] - 1
    neles = nx * ny
    x = _round10(xn).ravel()          # i-major, j fastest == node numbering
    y = _round10(yn).ravel()
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    nid = lambda i, j: i * (ny + 1) + j + 1
    eid = lambda i, j: i * ny + j + 1
    nod = np.stack([nid(ii, jj), nid(ii + 1, jj), nid(ii + 1, jj + 1), nid(ii, jj + 1)]).astype(np.int32)
    nc = np.stack([np.where(jj > 0, eid(ii, jj - 1), -1),
                   np.where(ii < nx - 1, eid(ii + 1, jj), -1),
                   np.where(jj < ny - 1, eid(ii, jj + 1), -1),
                   np.where(ii > 0, eid(ii - 1, jj), -1)]).astype(np.int32)
    side_type = {1: bc["bottom"], 2: bc["right"], 3: bc["top"], 4: bc["left"]}
    parents, ghosts, types, sides = [], [], [], []
    gid = neles
    bnd = np.nonzero((nc < 0).any(axis=0))[0]
    for e in bnd:
        for k in range(4):
            if nc[k, e] < 0:
                gid += 1
                nc[k, e] = gid
                parents.append(e + 1)
                ghosts.append(gid)
                types.append(side_type[k + 1])
                sides.append(k + 1)
    return Mesh(x, y, nod, nc, np.array(parents, np.int32), np.array(ghosts, np.int32),
                np.array(types, np.int32), np.array(sides, np.int32))


def _jitter(xn, yn, frac, rng):
    """Move interior nodes by +-frac of the smaller adjacent spacing, per direction."""
    if frac <= 0.0:
        return xn, yn
    xn, yn = xn.copy(), yn.copy()
    dxl = xn[1:-1, 1:-1] - xn[:-2, 1:-1]
    dxr = xn[2:, 1:-1] - xn[1:-1, 1:-1]
    dyl = yn[1:-1, 1:-1] - yn[1:-1, :-2]
    dyr = yn[1:-1, 2:] - yn[1:-1, 1:-1]
    hx = np.minimum(np.abs(dxl), np.abs(dxr))
    hy = np.minimum(np.abs(dyl), np.abs(dyr))
    shape = hx.shape
    xn[1:-1, 1:-1] += frac * hx * rng.uniform(-1.0, 1.0, size=shape)
    yn[1:-1, 1:-1] += frac * hy * rng.uniform(-1.0, 1.0, size=shape)
    return xn, yn


def channel(nx=97, ny=49, length=500.0, height=50.0, h1=0.01, bc=None, jitter=0.0,
            seed=DEFAULT_SEED) -> Mesh:
    """Planar channel / flat plate, wall-stretched in y (first cell h1 mm)."""
    bc = bc or {"left": 7, "right": 3, "bottom": 5, "top": 4}
    xs = np.array([length * i / nx for i in range(nx + 1)])
    ys = stretched(h1, height, ny)
    xn = np.repeat(xs[:, None], ny + 1, axis=1)
    yn = np.repeat(ys[None, :], nx + 1, axis=0)
    xn, yn = _jitter(xn, yn, jitter, np.random.default_rng(seed))
    return structured_mesh(xn, yn, bc)


def cosine_duct_radius(x: float, r_in=50.0, r_throat=25.0, x_throat=150.0, r_exit=40.0,
                       x_exit=400.0) -> float:
    """Half-cosine converging then diverging wall radius (mm)."""
    if x <= x_throat:
        return r_throat + (r_in - r_throat) * 0.5 * (1.0 + math.cos(math.pi * x / x_throat))
    s = (x - x_throat) / (x_exit - x_throat)
    return r_throat + (r_exit - r_throat) * 0.5 * (1.0 - math.cos(math.pi * s))


def hicks_henne(x: float, amplitudes, centres=None, length: float = 400.0,
                width: float = 3.0) -> float:
    """Sum of Hicks-Henne bumps a_k sin^width(pi t^(ln 0.5 / ln c_k)), t = x / length,
    each peaking (value a_k) at x = centre_k; centres default to len(amplitudes)
    stations evenly spaced over 50..350 mm.  Zero at both ends."""
    amplitudes = list(amplitudes)
    if not amplitudes:
        return 0.0
    if centres is None:
        k = len(amplitudes)
        centres = [50.0 + 300.0 * i / (k - 1) for i in range(k)] if k > 1 else [200.0]
    t = min(max(x / length, 0.0), 1.0)
    if t <= 0.0 or t >= 1.0:
        return 0.0
    out = 0.0
    for a, c in zip(amplitudes, centres):
        e = math.log(0.5) / math.log(c / length)
        out += a * math.sin(math.pi * t ** e) ** width
    return out


def duct(nx=385, ny=129, h1=0.005, jitter=0.01, seed=DEFAULT_SEED, bc=None,
         x_exit=400.0, radius=None, bumps=None) -> Mesh:
    """Axisymmetric converging-diverging duct; axis at y=0, wall on top.

    Radial stretching starts from a first cell h1 (mm) at the wall.  bumps:
    Hicks-Henne amplitudes (mm) added to the wall radius (centres over 50..350 mm).
    """
    bc = bc or {"left": 2, "right": 3, "bottom": 5, "top": 5}
    base = radius or (lambda xx: cosine_duct_radius(xx, x_exit=x_exit))
    radius = base if not bumps else (
        lambda xx: base(xx) + hicks_henne(xx, bumps, length=x_exit))
    xs = np.array([x_exit * i / nx for i in range(nx + 1)])
    xn = np.repeat(xs[:, None], ny + 1, axis=1)
    yn = np.zeros((nx + 1, ny + 1))
    for i in range(nx + 1):
        r = radius(xs[i])
        s = stretched(h1, r, ny)            # from the wall inwards
        yn[i, :] = r - s[::-1]
        yn[i, 0] = 0.0
    xn, yn = _jitter(xn, yn, jitter, np.random.default_rng(seed))
    return structured_mesh(xn, yn, bc)


def ramp(nx=2049, ny=513, length=1000.0, corner=400.0, angle_deg=15.0, height=300.0,
         h1=0.002, jitter=0.005, seed=DEFAULT_SEED, bc=None) -> Mesh:
    """2-D compression ramp: flat wall to x=corner, then a ramp; flat top at y=height."""
    bc = bc or {"left": 3, "right": 3, "bottom": 5, "top": 3}
    tan_a = math.tan(math.radians(angle_deg))
    xs = np.array([length * i / nx for i in range(nx + 1)])
    xn = np.repeat(xs[:, None], ny + 1, axis=1)
    yn = np.zeros((nx + 1, ny + 1))
    for i in range(nx + 1):
        yw = max(0.0, xs[i] - corner) * tan_a
        s = stretched(h1, height - yw, ny)
        yn[i, :] = yw + s
    xn, yn = _jitter(xn, yn, jitter, np.random.default_rng(seed))
    return structured_mesh(xn, yn, bc)


# --- named workloads -----------------------------------------------------------------------

def flow_channel(**kw) -> Flow:
    """Subsonic channel / flat plate flow (97x49 channel case)."""
    f = Flow(gama=1.4, amol=28.96, local=1, niter=2000, istart=0, cfl=0.25, prandtl=0.72,
             prandtlt=0.9, p0=1.2e5, t0=298.0, pamb=-1.0, amachinit=0.3, rtred=1e-5,
             npseudo=1, iaxis=0)
    return _with(f, kw)


def flow_duct(**kw) -> Flow:
    """Axisymmetric converging-diverging duct flow (385x129 duct case)."""
    f = Flow(gama=1.4, amol=28.96, local=1, niter=20000, istart=0, cfl=0.25, prandtl=0.72,
             prandtlt=0.9, p0=8.735e5, t0=298.0, pamb=-1.0, amachinit=0.01, rtred=1e-5,
             npseudo=1, iaxis=1)
    return _with(f, kw)


def flow_ramp(**kw) -> Flow:
    """Mach 2.5 compression-ramp flow (2049x513 ramp case)."""
    f = Flow(gama=1.4, amol=28.96, local=1, niter=20000, istart=0, cfl=0.25, prandtl=0.72,
             prandtlt=0.9, p0=1.7e6, t0=300.0, pamb=-1.0, amachinit=2.5, rtred=1e-5,
             npseudo=1, iaxis=0)
    return _with(f, kw)


def bl_inlet_params(flow: Flow, deltabl_m: float = 0.005) -> Flow:
    """Defines the BC type-7 inputs the reference leaves unset.

    PINF = static pressure of the initial Mach number, AMINF = AMACHINIT,
    DELTABL = boundary-layer thickness in metres (cell centroids are in m).
    """
    g = flow.gama
    dyn = 1.0 + (g - 1.0) * 0.5 * flow.amachinit * flow.amachinit
    return _with(flow, {"pinf": flow.p0 / dyn ** (g / (g - 1.0)), "aminf": flow.amachinit,
                        "deltabl": deltabl_m})


def sweep_case(i: int, seed: int = DEFAULT_SEED, nx_range=(65, 1025), ny_range=(33, 257)):
    """Case i of the batched parameter sweep, drawn from numpy's
    default_rng([seed, i]): family (the 97x49 channel's or the 2049x513 ramp's
    geometry, boundary types and flow) with equal odds, Nx and Ny log-uniform
    over nx_range / ny_range, AMACHINIT uniform in [0.2, 2.0], P0 uniform in
    [1e5, 2e6] Pa, T0 uniform in [250, 400] K; CFL 0.25, 5000 iterations.
    Channel cases get the BC type-7 inputs of bl_inlet_params.  A ramp case
    keeps the 2049-column ramp's node displacement in x (0.5 % of its spacing):
    0.5 % of a coarser spacing would move the first node row across the sloped
    wall, whose first cell is 0.002 mm tall, and such grids diverge at once in
    the reference's arithmetic.  Returns (family, Mesh, Flow)."""
    rng = np.random.default_rng([seed, i])
    family = "channel" if rng.random() < 0.5 else "ramp"
    nx = int(round(math.exp(rng.uniform(math.log(nx_range[0]), math.log(nx_range[1])))))
    ny = int(round(math.exp(rng.uniform(math.log(ny_range[0]), math.log(ny_range[1])))))
    mach = float(rng.uniform(0.2, 2.0))
    p0 = float(rng.uniform(1e5, 2e6))
    t0 = float(rng.uniform(250.0, 400.0))
    case_seed = int(rng.integers(0, 2 ** 31 - 1))
    if family == "channel":
        mesh = channel(nx=nx, ny=ny, seed=case_seed)
        flow = bl_inlet_params(flow_channel(amachinit=mach, p0=p0, t0=t0))
    else:
        mesh = ramp(nx=nx, ny=ny, seed=case_seed, jitter=0.005 * min(1.0, nx / 2049))
        flow = flow_ramp(amachinit=mach, p0=p0, t0=t0)
    return family, mesh, _with(flow, {"niter": 5000})


def sweep_cases(ncases: int = 64, seed: int = DEFAULT_SEED, nx_range=(65, 1025),
                ny_range=(33, 257)):
    """The batched parameter sweep: sweep_case(i) for i < ncases."""
    return [sweep_case(i, seed, nx_range, ny_range) for i in range(ncases)]


def _with(f: Flow, kw: dict) -> Flow:
    import dataclasses
    return dataclasses.replace(f, **kw)


WORKLOADS = {
    # name: (mesh factory, flow factory)
    "channel_97x49": (lambda seed=DEFAULT_SEED: channel(seed=seed),
                      lambda: bl_inlet_params(flow_channel())),
    "duct_385x129": (lambda seed=DEFAULT_SEED: duct(seed=seed), flow_duct),
    "ramp_2049x513": (lambda seed=DEFAULT_SEED: ramp(seed=seed), flow_ramp),
    # the same ramp with the 1/7-power boundary-layer inlet (type 7: p, T and u
    # fixed at PINF, TINF and the Mach 2.5 profile) on the left instead of
    # extrapolation (type 3).  Extrapolating every variable at a supersonic
    # inflow fixes none of them, and the type-3 ramp's inflow column drifts until
    # the run goes non-finite (iteration 3709 in FP64, in this library and in the
    # oracle alike); the type-7 ramp stays finite over its 20000 iterations.
    "ramp_2049x513_t7": (lambda seed=DEFAULT_SEED: ramp(seed=seed, bc={"left": 7, "right": 3,
                                                                        "bottom": 5, "top": 3}),
                         lambda: bl_inlet_params(flow_ramp())),
    # the adjoint configuration's grid: the duct at 193x65 with 8 Hicks-Henne bumps
    # (amplitudes 0 here; a design study sets them).  Its nodes move as far in x as
    # the 385-column duct's (1 % of that spacing): 1 % of the coarser spacing moves
    # near-wall nodes by four first-cell heights, and the grid diverges at once.
    "duct_193x65": (lambda seed=DEFAULT_SEED: duct(nx=193, ny=65, seed=seed, bumps=[0.0] * 8,
                                                   jitter=0.01 * 193 / 385),
                    flow_duct),
}
