"""The host-side mesh preparation (csrc/mesh_prep.cpp: validation, renumbering,
ghost tables, node-gather lists, BC type 6 port columns, the wave node tables)
under AddressSanitizer
and UndefinedBehaviorSanitizer, on the CPU: a g++ build of it with a small
driver (tests/native/mesh_prep_check.cpp) prepares fixtures and generated grids
in every ordering, and malformed grids, which it must reject with an error code
rather than read or write out of bounds."""
import copy
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG, REPO


def write_meshes(path, meshes):
    with open(path, "wb") as fh:
        np.array([len(meshes)], np.int32).tofile(fh)
        for m in meshes:
            np.array([len(m.x_mm), m.nod.shape[1], len(m.parent)], np.int32).tofile(fh)
            np.asarray(m.x_mm, np.float64).tofile(fh)
            np.asarray(m.y_mm, np.float64).tofile(fh)
            for a in (m.nod, m.nc, m.parent, m.ghost, m.btype, m.side):
                np.ascontiguousarray(a, np.int32).tofile(fh)


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("msan") / "mesh_prep_check")
    subprocess.check_call(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined",
                           "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
                           "-I" + os.path.join(REPO, "include"), "-I" + os.path.join(PKG, "csrc"),
                           os.path.join(REPO, "tests", "native", "mesh_prep_check.cpp"),
                           os.path.join(PKG, "csrc", "mesh_prep.cpp"), "-o", exe])
    return exe


def run(checker, tmp_path, meshes):
    path = str(tmp_path / "meshes.bin")
    write_meshes(path, meshes)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([checker, path], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    rows = []
    for ln in r.stdout.strip().splitlines():
        t = ln.split()
        rows.append({t[i]: int(t[i + 1]) for i in range(0, len(t), 2)})
    return rows


def test_valid_grids_prepare_cleanly(checker, tmp_path, golden):
    from tdt import gridgen
    meshes = [golden(n).mesh for n in ("t2d_chan", "t7_plate", "t6_port", "t6_port_axi", "case1",
                                       "taxi_duct")]
    meshes += [gridgen.channel(nx=16, ny=8), gridgen.WORKLOADS["duct_193x65"][0](),
               gridgen.ramp(nx=64, ny=32, jitter=0.005 * 64 / 2049)]
    rows = run(checker, tmp_path, meshes)
    assert len(rows) == 4 * len(meshes)
    for k, m in enumerate(meshes):
        mine = [r for r in rows if r["mesh"] == k]
        assert all(r["code"] == 0 for r in mine), mine
        assert {(r["n"], r["g"], r["m"]) for r in mine} == {(m.nod.shape[1], len(m.parent),
                                                             len(m.x_mm))}
        # RCM is kept only where it narrows the bandwidth of grid.in's numbering
        by = {r["order"]: r for r in mine}
        assert by[1]["bw"] <= by[0]["bw"] or by[1]["used"] == 0
        # the wave node tables (LDS node copies of FP64 k_dv): every cell's four nodes at
        # its slots of its wave's sorted, padded list, at most four 64-lane rounds
        assert all(r["nlok"] == 1 and 64 <= r["nlcap"] <= 256 for r in mine), mine


def _bad(m, what):
    m = copy.deepcopy(m)
    if what == "ghost_collides":
        m.ghost[0] = 3
    elif what == "type_unknown":
        m.btype[0] = 9
    elif what == "side_wrong":
        m.side[0] = 3
    elif what == "nc_dangling":
        m.nc[1, 0] = 10 ** 6
    elif what == "nc_negative":
        m.nc[2, 5] = -7
    elif what == "node_zero":
        m.nod[0, 0] = 0
    elif what == "node_past_end":
        m.nod[3, 4] = len(m.x_mm) + 1
    elif what == "parent_out_of_range":
        m.parent[1] = m.nod.shape[1] + 5
    elif what == "parent_zero":
        m.parent[2] = 0
    elif what == "side_zero":
        m.side[1] = 0
    elif what == "ghost_negative":
        m.ghost[0] = -1
    return m


BAD = ["ghost_collides", "type_unknown", "side_wrong", "nc_dangling", "nc_negative", "node_zero",
       "node_past_end", "parent_out_of_range", "parent_zero", "side_zero", "ghost_negative"]


def test_malformed_grids_are_rejected_without_memory_errors(checker, tmp_path, golden):
    base = golden("t2d_chan").mesh
    rows = run(checker, tmp_path, [_bad(base, w) for w in BAD])
    for k, what in enumerate(BAD):
        mine = [r for r in rows if r["mesh"] == k]
        assert mine and all(r["code"] < 0 for r in mine), (what, mine)
