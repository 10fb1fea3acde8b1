"""Host-side tests that need no GPU: the C ABI library, case I/O, grid
generation and mesh validation."""
import dataclasses

import numpy as np
import pytest

import tdt
from tdt import caseio, gridgen, lib


def test_library_exports_every_declared_symbol():
    L = tdt.load()
    syms = lib.header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert tdt.version().startswith("tdt")


def test_case_files_round_trip(tmp_path, golden):
    g = golden("t2d_ramp")
    caseio.write_case(str(tmp_path), g.flow, g.mesh)
    flow, mesh = caseio.load_case(str(tmp_path))
    assert dataclasses.replace(flow, pinf=g.flow.pinf, aminf=g.flow.aminf,
                               deltabl=g.flow.deltabl) == g.flow
    for k, v in g.mesh.arrays().items():
        assert np.array_equal(getattr(mesh, k), v), k
    c = g["final"]
    caseio.write_restart(str(tmp_path / "restart.in"), c, 1e-5, 1, 1.2345678901234567e-5, 30)
    c2, rtred, npseudo, rinit, niterold = caseio.read_restart(str(tmp_path / "restart.in"),
                                                              g.mesh.neles)
    assert np.array_equal(c, c2)
    assert (rtred, npseudo, rinit, niterold) == (1e-5, 1, 1.2345678901234567e-5, 30)


def test_truncated_restart_is_rejected(tmp_path, golden):
    g = golden("t2d_chan")
    caseio.write_restart(str(tmp_path / "r.in"), g["final"], 1e-5, 1, 1e-5, 30)
    text = open(tmp_path / "r.in").read().split("\n")
    open(tmp_path / "r.in", "w").write("\n".join(text[:10]))
    with pytest.raises(ValueError):
        caseio.read_restart(str(tmp_path / "r.in"), g.mesh.neles)


def test_duct_generator_reproduces_fixture(golden):
    g = golden("duct_385x129")
    m = gridgen.duct()
    for k, v in g.mesh.arrays().items():
        assert np.array_equal(getattr(m, k), v), k


def test_structured_numbering_matches_example_case_convention():
    m = gridgen.channel(nx=4, ny=3, bc={"left": 2, "right": 3, "bottom": 4, "top": 5})
    # cell 1: nodes (0,0),(1,0),(1,1),(0,1) -> 1, ny+2, ny+3, 2
    assert list(m.nod[:, 0]) == [1, 5, 6, 2]
    # NC2 of cell 1 is the cell to the right (id ny+1), NC3 the one above (id 2)
    assert m.nc[1, 0] == 4 and m.nc[2, 0] == 2
    # first ghosts: cell 1 bottom (side 1, type 4), then cell 1 left (side 4, type 2)
    assert list(zip(m.parent[:2], m.ghost[:2], m.btype[:2], m.side[:2])) == [
        (1, 13, 4, 1), (1, 14, 2, 4)]
    assert m.nghosts == 2 * (4 + 3)


def test_geometric_stretching():
    q = gridgen.geometric_ratio(0.01, 50.0, 49)
    assert abs(q - 1.143683) < 1e-6
    s = gridgen.stretched(0.01, 50.0, 49)
    assert s[0] == 0.0 and s[-1] == 50.0 and abs(s[1] - 0.01) < 1e-12


def _create(mesh, flow):
    return tdt.Solver(mesh, flow)


@pytest.mark.parametrize("mutate,code", [
    (lambda m: m.ghost.__setitem__(0, 3), lib.TDT_ERR_MESH),          # collides with a cell id
    (lambda m: m.btype.__setitem__(0, 1), lib.TDT_ERR_MESH),          # unknown type
    (lambda m: m.btype.__setitem__(0, 9), lib.TDT_ERR_MESH),          # unknown type
    (lambda m: m.side.__setitem__(0, 3), lib.TDT_ERR_MESH),           # wrong face
    (lambda m: m.nc.__setitem__((1, 0), 10**6), lib.TDT_ERR_MESH),    # dangling neighbour
    (lambda m: m.nod.__setitem__((0, 0), 0), lib.TDT_ERR_MESH),       # node id out of range
])
def test_mesh_validation(golden, mutate, code):
    g = golden("t2d_chan")
    m = caseio.Mesh.from_arrays({k: v.copy() for k, v in g.mesh.arrays().items()})
    mutate(m)
    with pytest.raises(tdt.TdtError) as ei:
        _create(m, g.flow)
    assert ei.value.code == code, str(ei.value)


def test_bc6_column_must_cross_a_cell():
    """A burning (type-6) face whose parent's opposite neighbour is a ghost (a
    one-cell-thick domain) has no port column to sum: refused."""
    m = gridgen.channel(nx=4, ny=1, bc={"left": 5, "right": 3, "bottom": 6, "top": 4})
    f = gridgen.flow_channel(tflame=2500.0, ancomb=0.35, acomb=1e-3, rhoprop=1700.0)
    with pytest.raises(tdt.TdtError) as ei:
        _create(m, f)
    assert ei.value.code == lib.TDT_ERR_UNSUPPORTED, str(ei.value)


def test_minus_one_neighbours_resolve_through_bc(golden):
    """NCk = -1 on a boundary face is accepted and resolved through bc.in
    (PARENT, SIDE); validation then succeeds and only the device check can fail."""
    g = golden("t2d_chan")
    m = caseio.Mesh.from_arrays({k: v.copy() for k, v in g.mesh.arrays().items()})
    m.nc[m.nc > m.neles] = -1
    try:
        s = _create(m, g.flow)
        s.close()
    except tdt.TdtError as e:
        assert e.code == lib.TDT_ERR_DEVICE, str(e)


def test_no_gpu_fails_loudly(golden):
    """Without a GPU the library refuses to run; there is no CPU fallback."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    g = golden("t2d_chan")
    with pytest.raises(tdt.TdtError) as ei:
        _create(g.mesh, g.flow)
    assert ei.value.code == lib.TDT_ERR_DEVICE


def test_hicks_henne_bumps_and_adjoint_grid():
    """Each Hicks-Henne bump peaks at its centre with its amplitude and vanishes
    at both ends; the 193x65 duct has the adjoint configuration's sizes and its
    wall moves by the bump amplitudes."""
    from tdt import gridgen
    for c in (50.0, 92.857, 350.0):
        assert abs(gridgen.hicks_henne(c, [0.7], [c]) - 0.7) < 1e-12
    assert gridgen.hicks_henne(0.0, [1.0] * 8) == 0.0
    assert gridgen.hicks_henne(400.0, [1.0] * 8) == 0.0
    m = gridgen.WORKLOADS["duct_193x65"][0]()
    assert (m.neles, m.nghosts, m.nods) == (12545, 516, 12804)
    amps = [0.3] * 8
    flat = gridgen.duct(nx=193, ny=65, jitter=0.0)
    bumped = gridgen.duct(nx=193, ny=65, jitter=0.0, bumps=amps)
    top = np.arange(65, flat.nods, 66)               # the wall node of each station
    x = flat.x_mm[top]
    want = np.array([gridgen.hicks_henne(v, amps) for v in x])
    assert np.allclose(bumped.y_mm[top] - flat.y_mm[top], want, atol=1e-9)


def _sweep(n, seed=5):
    return [(m, f) for _, m, f in gridgen.sweep_cases(n, seed=seed, nx_range=(12, 20),
                                                      ny_range=(8, 12))]


@pytest.mark.parametrize("mutate,code", [
    (None, lib.TDT_ERR_DEVICE),                                         # valid: device check
    (lambda c: dataclasses.replace(c, local=0), lib.TDT_ERR_UNSUPPORTED),
    (lambda c: dataclasses.replace(c, iaxis=1), lib.TDT_ERR_UNSUPPORTED),
    (lambda c: dataclasses.replace(c, gama=1.3), lib.TDT_ERR_UNSUPPORTED),
    (lambda c: dataclasses.replace(c, cfl=0.0), lib.TDT_ERR_ARG),
])
def test_batch_validation(mutate, code):
    """A batch checks every case as tdt_create would, and that the cases share
    the gas and model constants, before it looks for a device."""
    cases = _sweep(3)
    if mutate:
        m, f = cases[1]
        cases[1] = (m, mutate(f))
    with pytest.raises(tdt.TdtError) as ei:
        tdt.Batch(cases)
    assert ei.value.code == code, str(ei.value)


def test_batch_names_the_bad_case():
    cases = _sweep(3)
    m, f = cases[2]
    m = caseio.Mesh.from_arrays({k: v.copy() for k, v in m.arrays().items()})
    m.nod[0, 0] = 0
    cases[2] = (m, f)
    with pytest.raises(tdt.TdtError) as ei:
        tdt.Batch(cases)
    assert ei.value.code == lib.TDT_ERR_MESH and "case 2" in str(ei.value)


def test_sweep_cases_are_reproducible_and_in_range():
    a = gridgen.sweep_cases(6, seed=11, nx_range=(65, 129), ny_range=(33, 65))
    b = gridgen.sweep_cases(6, seed=11, nx_range=(65, 129), ny_range=(33, 65))
    for (fa, ma, qa), (fb, mb, qb) in zip(a, b):
        assert fa == fb and qa == qb
        for k, v in ma.arrays().items():
            assert np.array_equal(v, getattr(mb, k)), k
    for fam, m, f in a:
        assert fam in ("channel", "ramp")
        assert 0.2 <= f.amachinit <= 2.0 and 1e5 <= f.p0 <= 2e6 and 250.0 <= f.t0 <= 400.0
        assert f.cfl == 0.25 and f.niter == 5000 and f.local == 1
        assert 65 * 33 <= m.neles <= 129 * 65
        if fam == "channel":
            assert f.deltabl > 0 and f.aminf == f.amachinit


def test_ramp_with_type7_inflow():
    """The type-7-inflow ramp is the 2049x513 ramp with only its left ghosts
    changed (type 3 -> 7), and its inlet values are the Mach 2.5 free stream:
    PINF = P0 / (1 + (g-1)/2 M^2)^(g/(g-1))."""
    from tdt import gridgen
    m3 = gridgen.WORKLOADS["ramp_2049x513"][0]()
    m7 = gridgen.WORKLOADS["ramp_2049x513_t7"][0]()
    f3 = gridgen.WORKLOADS["ramp_2049x513"][1]()
    f7 = gridgen.WORKLOADS["ramp_2049x513_t7"][1]()
    assert np.array_equal(m3.x_mm, m7.x_mm) and np.array_equal(m3.y_mm, m7.y_mm)
    assert np.array_equal(m3.nc, m7.nc) and np.array_equal(m3.parent, m7.parent)
    left = m3.side == 4
    assert left.sum() == 513
    assert np.all(m3.btype[left] == 3) and np.all(m7.btype[left] == 7)
    assert np.array_equal(m3.btype[~left], m7.btype[~left])
    assert f7.aminf == 2.5 and f7.deltabl > 0.0
    assert abs(f7.pinf - 1.7e6 / 2.25 ** 3.5) < 1e-6 * f7.pinf
    assert (f7.p0, f7.t0, f7.amachinit, f7.niter) == (f3.p0, f3.t0, f3.amachinit, f3.niter)


def test_wall_faces_of_each_workload():
    """No-slip faces for the wall shear comparison: the channel's and the ramp's
    whole bottom (planar walls at y = 0 included), the duct's upper wall only
    (its type-5 axis at y = 0 is not a wall)."""
    from tdt import analysis, gridgen
    for name, want in (("channel_97x49", 97), ("ramp_2049x513", 2049), ("duct_385x129", 385)):
        mk_mesh, mk_flow = gridgen.WORKLOADS[name]
        m, f = mk_mesh(), mk_flow()
        faces = analysis.wall_faces(m, f.iaxis)
        assert len(faces) == want, name
        ymid = 0.5 * (m.y_mm[faces[:, 1]] + m.y_mm[faces[:, 2]])
        if f.iaxis:
            assert np.all(ymid > 1.0), name          # the upper wall, radius >= 25 mm
        else:
            assert np.all(ymid <= 600.0 * np.tan(np.radians(15.0)) * 1.0001), name  # ramp top
            assert np.sum(ymid == 0.0) >= 97 * (name == "channel_97x49") + 819 * (name != "channel_97x49")


def test_duct_objective_gradients():
    """J = axial wall-pressure force + outlet mass flow of the 193x65 duct:
    dJ/dc and the explicit dJ/dY against central differences, and the bump
    node derivatives against a direct evaluation of the generator."""
    import dataclasses
    from tdt import gridgen, objective
    mesh, flow = gridgen.WORKLOADS["duct_193x65"][0](), gridgen.WORKLOADS["duct_193x65"][1]()
    rng = np.random.default_rng(2)
    n = mesh.neles
    c = np.stack([np.full(n, 1.2), np.full(n, 60.0), np.zeros(n), np.full(n, 2.6e5),
                  np.full(n, 1.7e-5)]) * (1.0 + 0.01 * rng.standard_normal((5, n)))
    j0, dc, dxy = objective.duct_objective(mesh, flow, c)
    assert len(objective.analysis.wall_faces(mesh, flow.iaxis)) == 193
    assert len(objective.outlet_faces(mesh)) == 65
    w = c * rng.standard_normal(c.shape)
    h = 1e-7
    fd = (objective.duct_objective(mesh, flow, c + h * w)[0]
          - objective.duct_objective(mesh, flow, c - h * w)[0]) / (2 * h)
    assert abs(fd - np.sum(dc * w)) <= 1e-7 * np.sum(np.abs(dc * w))
    wy = rng.standard_normal(mesh.nods) * 1e-2
    jp = objective.duct_objective(dataclasses.replace(mesh, y_mm=mesh.y_mm + h * wy), flow, c)[0]
    jm = objective.duct_objective(dataclasses.replace(mesh, y_mm=mesh.y_mm - h * wy), flow, c)[0]
    assert abs((jp - jm) / (2 * h) - np.sum(dxy[1] * wy)) <= 1e-6 * np.sum(np.abs(dxy[1] * wy))
    assert np.all(dxy[0] == 0.0)
    kw = dict(nx=193, ny=65, jitter=0.01 * 193 / 385)
    dn = objective.bump_node_derivatives(8, **kw)
    assert dn.shape == (8, 2, mesh.nods) and np.all(dn[:, 0] == 0.0)
    # the wall node under bump 3's centre moves by 1 per mm of amplitude
    top = np.arange(65, mesh.nods, 66)
    x3 = 50.0 + 300.0 * 3 / 7
    col = np.argmin(np.abs(mesh.x_mm[top] - x3))
    assert abs(dn[3, 1, top[col]] - gridgen.hicks_henne(mesh.x_mm[top[col]], [1.0], [x3])) < 1e-6
