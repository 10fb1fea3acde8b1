"""Routine-level GPU parity: each libtdt entry point against the reference dumps.

Every routine is fed the reference's own inputs (injected with set_array), so
a mismatch is attributable to that routine alone.  Routines whose arithmetic
involves no libm call must agree BIT FOR BIT (same operations, same order, no
contraction).  Where the reference calls pow / tanh (boundary types 2, 6 and 7,
viscous fluxes, source terms), the device math library may differ from glibc
by an ulp; those arrays are held to 1e-13 of the array's largest magnitude.
"""
import numpy as np
import pytest

from conftest import DUMP_CASES

pytestmark = pytest.mark.gpu

ALPHAS = (0.25, 1.0 / 3.0, 0.5, 1.0)
ULP_TOL = 1e-13


def exact(name, got, want):
    got, want = np.asarray(got), np.asarray(want)
    assert got.shape == want.shape, f"{name}: shape {got.shape} != {want.shape}"
    bad = ~((got == want) | (np.isnan(got) & np.isnan(want)))
    if bad.any():
        i = int(np.argmax(bad))
        raise AssertionError(f"{name}: {int(bad.sum())}/{got.size} differ, first at {i}: "
                             f"{got[i]!r} vs {want[i]!r}")


def close(name, got, want, tol=ULP_TOL):
    got, want = np.asarray(got), np.asarray(want)
    assert got.shape == want.shape
    scale = max(np.max(np.abs(want)), 1e-300)
    err = np.max(np.abs(got - want)) / scale
    assert err <= tol, f"{name}: max |diff| / max |ref| = {err:.3e} > {tol:.1e}"


def ghost_ids(g):
    return g.mesh.ghost - 1


@pytest.mark.parametrize("ordering", ["input", "rcm"])
@pytest.mark.parametrize("case", DUMP_CASES)
def test_routines_match_reference(golden, case, ordering):
    import tdt
    g = golden(case)
    ne = g.mesh.neles
    with tdt.Solver(g.mesh, g.flow, "fp64", ordering) as s:
        rinit = s.cold_start()
        exact("RTOL", s.get_array("RTOL"), g.dump("RTOL"))
        nmax = s.nmax
        assert nmax == g.dump("XCC").shape[0]
        for nm in ("XCC", "YCC", "VOL"):
            exact(nm, s.get_array(nm), g.dump(nm))
        exact("DL", s.get_array("DL"), g.dump("DL"))
        for k in range(1, 5):
            for ax in "XY":
                exact(f"SC{k}{ax}", s.get_array(f"SC{k}{ax}"), g.dump(f"SC{k}{ax}"))
        c0 = np.stack([g.dump(f"C{m}") for m in range(1, 6)])
        s.set_state(c0, rinit)
        # time step (turbnstwod.f90:97-114)
        s.timestep()
        for nm in "UVTP":
            exact("TS_" + nm, s.get_array(nm)[:ne], g.dump("TS_" + nm))
        exact("TS_DT", s.get_array("DT")[:ne], g.dump("TS_DT"))
        exact("DTMIN", s.get_array("DTMIN"), g.dump("DTMIN"))
        # dissipation (solver_state.f90:70-225)
        s.dissipation()
        exact("DS_P", s.get_array("P"), g.dump("DS_P"))
        exact("DS_DT", s.get_array("DT"), g.dump("DS_DT"))
        exact("ANU", s.get_array("ANU"), g.dump("ANU"))
        exact("ANURT", s.get_array("ANURT"), g.dump("ANURT"))
        for m in range(1, 6):
            exact(f"D2Q{m}", s.get_array(f"D2Q{m}"), g.dump(f"D2Q{m}"))
            exact(f"D24{m}", s.get_array(f"D24{m}"), g.dump(f"D24{m}"))
        gid = ghost_ids(g)
        pow_types = np.isin(g.mesh.btype, (2, 6, 7))
        for st in range(1, 5):
            S = f"S{st}"
            if st > 1:   # stage input = the reference's previous stage output
                for m in range(1, 6):
                    full = np.zeros(nmax)
                    full[:ne] = g.dump(f"S{st - 1}_Q{m}")
                    s.set_array(f"C{m}", full)
            s.primitives()
            for nm in "UVTP":
                exact(f"{S}_{nm}(cells)", s.get_array(nm)[:ne], g.dump(f"{S}_{nm}")[:ne])
            for p in "FG":
                for m in range(1, 6):
                    exact(f"{S}_{p}{m}(cells)", s.get_array(f"{p}{m}")[:ne], g.dump(f"{S}_{p}{m}")[:ne])
            s.apply_boundary()
            names = list("UVTP") + [f"{p}{m}" for p in "CFG" for m in range(1, 6)]
            for nm in names:
                got = s.get_array(nm)[gid]
                want = g.dump(f"{S}_{nm}")[gid]
                exact(f"{S}_{nm}(ghosts 3/4/5)", got[~pow_types], want[~pow_types])
                if pow_types.any():
                    close(f"{S}_{nm}(ghosts 2/6/7)", got[pow_types], want[pow_types])
            for nm in names:       # isolate what follows from libm differences
                s.set_array(nm, g.dump(f"{S}_{nm}"))
            if st == 1:
                s.viscous()
                for nm in ("UN", "VN", "TN", "RHON", "RN"):
                    exact(nm, s.get_array(nm), g.dump(nm))
                for nm in ("VISU", "VISV", "VIST", "VISR"):
                    close(nm, s.get_array(nm), g.dump(nm))
                    s.set_array(nm, g.dump(nm))
            s.source_terms()
            close(f"{S}_SRC5", s.get_array("SOURCE5"), g.dump(f"{S}_SRC5"))
            close(f"{S}_SRC3", s.get_array("SOURCE3"), g.dump(f"{S}_SRC3"))
            s.set_array("SOURCE5", g.dump(f"{S}_SRC5"))
            s.set_array("SOURCE3", g.dump(f"{S}_SRC3"))
            s.residual()
            for m in range(1, 6):
                exact(f"{S}_R{m}", s.get_array(f"R{m}"), g.dump(f"{S}_R{m}"))
            s.update(ALPHAS[st - 1])
            for m in range(1, 6):
                exact(f"{S}_Q{m}", s.get_array(f"C{m}")[:ne], g.dump(f"{S}_Q{m}"))
        l2, linf = s.monitor()
        exact("LINF", linf, g.dump("LINF"))
        np.testing.assert_allclose(l2, g.dump("L2"), rtol=1e-13)
        exact("COLD", np.stack([s.get_array(f"COLD{m}") for m in range(1, 6)]), g["final"])


@pytest.mark.parametrize("case", ["t2d_chan", "taxi_duct"])
def test_routine_chain_without_injection(golden, case):
    """The routines chained as the reference calls them, no injected inputs:
    one full iteration stays within 1e-12 of the reference."""
    import tdt
    g = golden(case)
    with tdt.Solver(g.mesh, g.flow, "fp64") as s:
        rinit = s.cold_start()
        s.set_state(np.stack([g.dump(f"C{m}") for m in range(1, 6)]), rinit)
        s.timestep()
        s.dissipation()
        for st in range(4):
            s.primitives()
            s.apply_boundary()
            if st == 0:
                s.viscous()
            s.source_terms()
            s.residual()
            s.update(ALPHAS[st])
        s.monitor()
        got = s.get_state()
        for m in range(5):
            close(f"C{m + 1}", got[m], g["final"][m], tol=1e-12)


def _postprocess_restated(mesh, flow, p, u, v, t):
    """solver_state.f90:230-283 in numpy: the same contributions in the same order."""
    rr = 8314.0 / flow.amol
    nn = mesh.nods
    acc = np.zeros((4, nn))
    cnt = np.zeros(nn)
    vals = np.stack([p, u, v, t])
    for i in range(mesh.neles):                   # cells by id, their four nodes
        for j in range(4):
            k = mesh.nod[j, i] - 1
            acc[:, k] = acc[:, k] + vals[:, i]
            cnt[k] += 1
    for g in range(mesh.nghosts):                 # then the ghosts in bc.in order
        ng, npar, ns = mesh.ghost[g] - 1, mesh.parent[g] - 1, mesh.side[g]
        n1, n2 = mesh.nod[ns - 1, npar] - 1, mesh.nod[ns % 4, npar] - 1
        for k in (n1, n2):
            acc[:, k] = acc[:, k] + vals[:, ng]
            cnt[k] += 1
    pn, un, vn, tn = acc / cnt
    return {"U": un, "V": vn, "P_bar": pn * 0.00001, "RHO": pn / (rr * tn),
            "MACH": np.sqrt(un * un + vn * vn) / np.sqrt(flow.gama * rr * tn), "T": tn}


@pytest.mark.parametrize("case", ["case1", "taxi_duct"])
def test_postprocess_matches_restatement(golden, case):
    """tdt_postprocess reproduces POSTPROCESS's node values bit for bit from the
    P, U, V, T arrays the routines leave (here after a few fused iterations and
    one routine-level stage)."""
    import tdt
    g = golden(case)
    with tdt.Solver(g.mesh, g.flow, "fp64") as s:
        s.cold_start()
        s.iterate(5)
        s.primitives()
        s.apply_boundary()
        got = s.postprocess()
        arrs = {k: s.get_array(k) for k in ("P", "U", "V", "T")}
    want = _postprocess_restated(g.mesh, g.flow, arrs["P"], arrs["U"], arrs["V"], arrs["T"])
    for k in want:
        assert np.array_equal(got[k], want[k]), (k, np.max(np.abs(got[k] - want[k])))
