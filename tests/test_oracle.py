"""Pins the C oracle (oracle/tdt_oracle.c) to the reference solver.

The fixtures were written by tests/golden/make_golden.py from the reference's
own modules (cross-checked against the reference executable).  The oracle
must reproduce every dumped array and the whole iteration history BIT FOR BIT.
"""
import numpy as np
import pytest

from conftest import DUMP_CASES, RUN_CASES
from oracle import Oracle

ALPHAS = (0.25, 1.0 / 3.0, 0.5, 1.0)


def _eq(name, got, want):
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, f"{name}: shape {got.shape} != {want.shape}"
    bad = ~((got == want) | (np.isnan(got) & np.isnan(want)))
    if bad.any():
        i = int(np.argmax(bad))
        raise AssertionError(f"{name}: {int(bad.sum())} mismatches, first at {i}: "
                             f"{got.ravel()[i]!r} != {want.ravel()[i]!r}")


@pytest.mark.parametrize("case", DUMP_CASES)
def test_oracle_routines_bit_exact(golden, case):
    g = golden(case)
    o = Oracle(g.mesh, g.flow)
    o.cold_start()
    l2, linf = o.iterate(g.nrun)
    _eq("hist_l2", l2, g["hist_l2"].reshape(-1, 5))
    _eq("hist_linf", linf, g["hist_linf"].reshape(-1, 5))
    ne = g.mesh.neles
    nmax = g.dump("XCC").shape[0]
    for nm in ("XCC", "YCC", "VOL"):
        _eq(nm, o.array(nm)[:nmax], g.dump(nm))
    _eq("DL", o.array("DL"), g.dump("DL"))
    for k in range(1, 5):
        for ax in "XY":
            _eq(f"SC{k}{ax}", o.array(f"SC{k}{ax}"), g.dump(f"SC{k}{ax}"))
    _eq("RTOL", o.scalar("RTOL"), g.dump("RTOL")[0])
    for m in range(1, 6):
        _eq(f"C{m}", o.array(f"C{m}")[:ne], g.dump(f"C{m}"))
    o.timestep()
    for nm in "UVTP":
        _eq("TS_" + nm, o.array(nm)[:ne], g.dump("TS_" + nm))
    _eq("TS_DT", o.array("DT")[:ne], g.dump("TS_DT"))
    _eq("DTMIN", o.scalar("DTMIN"), g.dump("DTMIN")[0])
    o.dissipation()
    _eq("DS_P", o.array("P")[:nmax], g.dump("DS_P"))
    _eq("DS_DT", o.array("DT")[:nmax], g.dump("DS_DT"))
    _eq("ANU", o.array("ANU")[:nmax], g.dump("ANU"))
    _eq("ANURT", o.array("ANURT")[:nmax], g.dump("ANURT"))
    for m in range(1, 6):
        _eq(f"D2Q{m}", o.array(f"D2Q{m}")[:nmax], g.dump(f"D2Q{m}"))
        _eq(f"D24{m}", o.array(f"D24{m}"), g.dump(f"D24{m}"))
    for s in range(1, 5):
        S = f"S{s}"
        o.primitives()
        o.apply_boundary()
        for nm in "UVTP":
            _eq(f"{S}_{nm}", o.array(nm)[:nmax], g.dump(f"{S}_{nm}"))
        for m in range(1, 6):
            _eq(f"{S}_C{m}", o.array(f"C{m}")[:nmax], g.dump(f"{S}_C{m}"))
            _eq(f"{S}_F{m}", o.array(f"F{m}")[:nmax], g.dump(f"{S}_F{m}"))
            _eq(f"{S}_G{m}", o.array(f"G{m}")[:nmax], g.dump(f"{S}_G{m}"))
        if s == 1:
            o.viscous()
            for nm in ("UN", "VN", "TN", "RHON", "RN"):
                _eq(nm, o.array(nm), g.dump(nm))
            for nm in ("VISU", "VISV", "VIST", "VISR"):
                _eq(nm, o.array(nm), g.dump(nm))
        o.source_terms()
        _eq(f"{S}_SRC3", o.array("SOURCE3"), g.dump(f"{S}_SRC3"))
        _eq(f"{S}_SRC5", o.array("SOURCE5"), g.dump(f"{S}_SRC5"))
        o.residual()
        for m in range(1, 6):
            _eq(f"{S}_R{m}", o.array(f"R{m}"), g.dump(f"{S}_R{m}"))
        o.update(ALPHAS[s - 1])
        for m in range(1, 6):
            _eq(f"{S}_Q{m}", o.array(f"C{m}")[:ne], g.dump(f"{S}_Q{m}"))
    l2, linf = o.monitor()
    _eq("L2", l2, g.dump("L2"))
    _eq("LINF", linf, g.dump("LINF"))
    _eq("final", o.state(), g["final"])


@pytest.mark.parametrize("case", RUN_CASES)
def test_oracle_history_bit_exact(golden, case):
    g = golden(case)
    o = Oracle(g.mesh, g.flow)
    o.cold_start()
    l2, linf = o.iterate(g.nrun)
    _eq("hist_l2", l2, g["hist_l2"].reshape(-1, 5))
    _eq("hist_linf", linf, g["hist_linf"].reshape(-1, 5))
    final = o.state()
    if g.has("final_idx"):
        _eq("final(sample)", final[:, g["final_idx"]], g["final"])
        np.testing.assert_allclose(np.sqrt((final * final).sum(axis=1)), g["final_norm2"],
                                   rtol=1e-15)
    else:
        _eq("final", final, g["final"])
