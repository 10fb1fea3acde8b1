"""Batched cases on the GPU (tdt_batch_*): every case of a batch must come out
exactly as it does in a context of its own.

The batch renumbers each case as tdt_create would and gives it its own cell
blocks, so in FP64 the states, the L2 and the Linf histories of each case are
bitwise those of the single-case runs (same kernels, same blocks, same fold
order).  The single-case runs are themselves pinned to the reference by
test_gpu_fused.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ITERS = 30


def small_sweep(n, seed=7):
    """Sweep cases (both families, random Mach/P0/T0) on small grids."""
    from tdt import gridgen
    return [(m, f) for _, m, f in gridgen.sweep_cases(n, seed=seed, nx_range=(12, 60),
                                                      ny_range=(8, 30))]


def planar_cases(golden):
    cases = small_sweep(5)
    for name in ("t2d_chan", "t7_plate", "case1"):
        g = golden(name)
        cases.append((g.mesh, g.flow))
    return cases


def axi_cases(golden):
    from tdt import gridgen
    cases = [(golden(n).mesh, golden(n).flow) for n in ("taxi_duct", "case1_axi")]
    cases.append((gridgen.duct(nx=40, ny=16, jitter=0.0), gridgen.flow_duct(amachinit=0.3)))
    return cases


def run_single(mesh, flow, precision, iters):
    import tdt
    with tdt.Solver(mesh, flow, precision) as s:
        s.cold_start()
        l2, linf, bad = s.iterate(iters)
        assert bad == 0
        return s.get_state(), l2, linf


SET_NAMES = {"merged": "merged", "split": "split", "18": "split+preload"}


def check_bitwise(cases, precision, kernels, monkeypatch, iters=ITERS):
    import tdt
    monkeypatch.setenv("TDT_KERNELS", kernels)
    with tdt.Batch(cases, precision) as b:
        b.cold_start()
        l2, linf, bad = b.iterate(iters)
        assert bad == 0
        assert b.kernel_set() == SET_NAMES[kernels]
        states = [b.get_state(k) for k in range(len(cases))]
    for k, (mesh, flow) in enumerate(cases):
        st, l2s, linfs = run_single(mesh, flow, precision, iters)
        assert np.array_equal(states[k], st), \
            f"case {k}: max diff {np.max(np.abs(states[k] - st))}"
        assert np.array_equal(l2[:, k], l2s), f"case {k}: L2 history"
        assert np.array_equal(linf[:, k], linfs), f"case {k}: Linf history"


@pytest.mark.parametrize("kernels", ["merged", "split", "18"])
def test_batch_equals_single_cases_planar(golden, kernels, monkeypatch):
    check_bitwise(planar_cases(golden), "fp64", kernels, monkeypatch)


def test_batch_equals_single_cases_axisymmetric(golden, monkeypatch):
    check_bitwise(axi_cases(golden), "fp64", "merged", monkeypatch)


def test_batch_with_a_boundary_last_case(monkeypatch):
    """A case of 263,680 cells is renumbered boundary-last inside the batch, as it
    would be alone; small cases around it keep RCM."""
    from tdt import gridgen
    cases = small_sweep(2, seed=3)
    cases.insert(1, (gridgen.ramp(nx=1030, ny=256), gridgen.flow_ramp()))
    check_bitwise(cases, "fp64", "split", monkeypatch, iters=8)


def test_batch_fp32_mixed_tracks_single_cases(golden, monkeypatch):
    """FP32-mixed: the batch and single-case runs use kernels compiled apart, so
    FMA contraction may differ; they agree to FP32 rounding."""
    import tdt
    monkeypatch.setenv("TDT_KERNELS", "merged")
    cases = planar_cases(golden)
    with tdt.Batch(cases, "fp32-mixed") as b:
        b.cold_start()
        b.iterate(ITERS)
        states = [b.get_state(k) for k in range(len(cases))]
    for k, (mesh, flow) in enumerate(cases):
        st, _, _ = run_single(mesh, flow, "fp32-mixed", ITERS)
        a = tdt.primitives_from_state(states[k], flow)
        r = tdt.primitives_from_state(st, flow)
        for q in ("rho", "p"):
            err = np.linalg.norm(a[q] - r[q]) / np.linalg.norm(r[q])
            assert err < 1e-6, (k, q, err)


def test_batch_restart_and_state_round_trip(golden, monkeypatch):
    """set_state / get_state per case: 20 + 20 iterations with the states written
    back in between equal 40 straight iterations, bit for bit."""
    import tdt
    monkeypatch.setenv("TDT_KERNELS", "merged")
    cases = planar_cases(golden)[:4]
    with tdt.Batch(cases) as a, tdt.Batch(cases) as b:
        ra = a.cold_start()
        b.cold_start()
        a.iterate(20)
        mid = [a.get_state(k) for k in range(len(cases))]
        for k in range(len(cases)):
            a.set_state(k, mid[k], ra[k])
            assert np.array_equal(a.get_state(k), mid[k])
        a.iterate(20)
        b.iterate(40)
        for k in range(len(cases)):
            assert np.array_equal(a.get_state(k), b.get_state(k)), k


def test_batch_refuses_single_case_calls(golden):
    import tdt
    from tdt import lib
    cases = planar_cases(golden)[:2]
    with tdt.Batch(cases) as b:
        b.cold_start()
        for call in (lambda: b.L.tdt_timestep(b.h), lambda: b.L.tdt_iterate(b.h, 1, None, None, None),
                     lambda: b.L.tdt_cold_start(b.h, None)):
            assert call() == lib.TDT_ERR_UNSUPPORTED
        buf = np.zeros(5 * max(b.neles))
        assert b.L.tdt_batch_get_state(b.h, len(cases), lib._d(buf)) == lib.TDT_ERR_ARG
        assert b.L.tdt_batch_get_state(b.h, -1, lib._d(buf)) == lib.TDT_ERR_ARG


def sweep_with_fixtures(golden):
    """The batched sweep at full size (gridgen.sweep_cases(64): 36 channel and 28
    ramp cases, 2,685,047 cells) followed by two reference fixtures: the 97x49
    channel as specified (type-7 inlet) and the reference's example case."""
    from tdt import gridgen
    cases = [(m, f) for _, m, f in gridgen.sweep_cases(64)]
    fixtures = ["channel_97x49", "case1"]
    for name in fixtures:
        g = golden(name)
        cases.append((g.mesh, g.flow))
    return cases, fixtures


@pytest.mark.parametrize("precision", ["fp64", "fp32-mixed"])
def test_full_sweep_batch(golden, precision, monkeypatch):
    """The 64-case sweep in one batch context, a few dozen iterations: FP64 every
    case bit for bit its single-case run (state, L2 and Linf histories);
    FP32-mixed every case within FP32 rounding of its single-case run (kernels
    compiled apart).  In FP64 the fixture cases' L2 histories inside the batch
    agree with the reference's own to 1e-10."""
    import tdt
    monkeypatch.setenv("TDT_KERNELS", "merged")
    cases, fixtures = sweep_with_fixtures(golden)
    with tdt.Batch(cases, precision) as b:
        b.cold_start()
        l2, linf, bad = b.iterate(ITERS)
        assert bad == 0
        states = [b.get_state(k) for k in range(len(cases))]
    worst = 0.0
    for k, (mesh, flow) in enumerate(cases):
        st, l2s, linfs = run_single(mesh, flow, precision, ITERS)
        if precision == "fp64":
            assert np.array_equal(states[k], st), \
                f"case {k}: max diff {np.max(np.abs(states[k] - st))}"
            assert np.array_equal(l2[:, k], l2s), f"case {k}: L2 history"
            assert np.array_equal(linf[:, k], linfs), f"case {k}: Linf history"
        else:
            a = tdt.primitives_from_state(states[k], flow)
            r = tdt.primitives_from_state(st, flow)
            for q in ("rho", "p"):
                err = np.linalg.norm(a[q] - r[q]) / np.linalg.norm(r[q])
                worst = max(worst, err)
                assert err < 1e-6, (k, q, err)
    if precision != "fp64":
        return   # the L2 of C - COLD is a difference: FP32 rounding shows at 1e-3 there
    for j, name in enumerate(fixtures):
        k = len(cases) - len(fixtures) + j
        ref = golden(name)["hist_l2"].reshape(-1, 5)[:ITERS]
        dev = float(np.max(np.abs(l2[:, k] - ref) / np.maximum(np.abs(ref), 1e-300)))
        print(f"{precision}: batch case {k} ({name}) L2 history vs reference: {dev:.2e}")
        assert dev <= 1e-10, (name, dev)
    print(f"{precision}: {len(cases)} cases, {sum(m.neles for m, _ in cases)} cells; "
          f"worst FP32-mixed batch/single deviation {worst:.2e}")
