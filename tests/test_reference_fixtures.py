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


DUCT_CONVERGED = os.path.join(GOLDEN, "duct_193x65_converged.npz")


@pytest.mark.skipif(not os.path.exists(DUCT_CONVERGED),
                    reason="duct_193x65_converged fixture not generated")
def test_reference_adjoint_duct_converges_at_cfl_half():
    """The 193x65 adjoint duct at CFL 0.5 reaches a steady state in the
    reference's own arithmetic: its L2 history drops below 1e-8 of the first
    iterations' near 700,000 iterations and below 1e-10 by 850,000, so the
    central differences of its J are those of the steady state.  The moved runs'
    J straddle the base run's (second-order differences small against the
    first-order ones: the differences resolve the derivative)."""
    with np.load(DUCT_CONVERGED) as z:
        every = z["hist_l2_every1000"]
        head = z["hist_l2_head"]
        nrun = int(z["nrun"])
        jb = float(z["J_base"])
        jp = {k: (float(z[f"J_{k}+"]), float(z[f"J_{k}-"])) for k in ("p0", "t0", "bump3")}
    assert nrun == 850000 and len(every) == nrun // 1000
    drop = every / head[:10].max(axis=0)
    print(f"reference L2 drop at 400k {drop[399].max():.2e}, 700k {drop[699].max():.2e}, "
          f"850k {drop[-1].max():.2e}")
    assert drop[709:].max() < 1e-8 and drop[-1].max() < 1e-10
    assert drop[399].max() > 1e-7            # not settled before 400,000 iterations
    for key, (a, b) in jp.items():
        first = abs(a - b) / 2
        second = abs((a + b) / 2 - jb)
        assert first > 0 and second < 1e-3 * first, (key, first, second)
