"""The C ABI driven through several contexts in one process with libtdt's host
code under AddressSanitizer (tests/native/two_ctx_check.cpp, built by
`make -C 2d-turbulent-navier-stokes_amd checks`): per precision the benchmark's
sequence of calls (cold start, a first iterate that times the kernel sets with
event-bracketed launches and captures the graphs, a run with the history read
back, an event-timed eager run, destroy), FP64 then FP32-mixed then FP64, so
that every context after the first follows the destruction of another.  Any
host-side out-of-bounds access or use after free in that sequence stops the
program with an AddressSanitizer report."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "build", "two_ctx_asan")


def test_contexts_in_sequence_under_asan():
    assert os.path.exists(EXE), "build it first: make -C 2d-turbulent-navier-stokes_amd checks"
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:halt_on_error=1")
    r = subprocess.run([EXE], capture_output=True, text=True, env=env, timeout=600)
    print(r.stdout)
    assert "AddressSanitizer" not in r.stderr, r.stderr[-6000:]
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "two_ctx: OK" in r.stdout
