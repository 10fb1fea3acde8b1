"""What the reference solver's own runs (tests/golden, written by
tests/golden/make_golden.py from the reference's sources) record about the
specified configurations, checked on the CPU: facts that the GPU tests then
hold this library to."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

RAMP_T3 = os.path.join(GOLDEN, "ramp_2049x513_t3.npz")


@pytest.mark.skipif(not os.path.exists(RAMP_T3), reason="ramp_2049x513_t3 fixture not generated")
def test_reference_ramp_type3_goes_nonfinite():
    """The 2049x513 ramp as specified (type-3 extrapolation at its Mach 2.5
    inflow, outflow and top) is unstable in the reference's own arithmetic:
    its per-iteration change grows for hundreds of iterations (L-infinity of
    the energy change 20x from iteration 3200 to 3700) and its state goes
    non-finite shortly after iteration 3700, from the cold start, with no
    input from this library."""
    with np.load(RAMP_T3) as z:
        l2 = z["hist_l2"].reshape(-1, 5)
        linf = z["hist_linf"].reshape(-1, 5)
    finite = np.isfinite(l2).all(axis=1)
    first_bad = int(np.argmin(finite)) + 1 if not finite.all() else 0
    print(f"reference: first non-finite iteration {first_bad}; Linf(C4 change) at 3200 "
          f"{linf[3199, 3]:.3e}, at 3700 {linf[3699, 3]:.3e}")
    assert 3600 < first_bad <= len(l2), first_bad
    assert finite[:first_bad - 1].all()
    assert not finite[first_bad - 1:].any()
    assert linf[3699, 3] > 20.0 * linf[3199, 3]
    # the early history settles (the instability is not a start-up transient)
    assert linf[2999, 3] < 0.2 * linf[0, 3]
