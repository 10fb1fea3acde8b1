"""The hot path's FP64 division helpers on the GPU against the compiler's IEEE
division (and numpy's): a reused refined reciprocal must give the same bits as
a / b for finite, normal operands whose exponents are within a few hundred of
each other, which covers every division of a physical state."""
import ctypes
import os

import numpy as np
import pytest

from conftest import REPO

pytestmark = pytest.mark.gpu

LIB = os.path.join(REPO, "tests", "native", "build", "libdivcheck.so")


def run(a, b):
    assert os.path.exists(LIB), "build it first: make -C 2d-turbulent-navier-stokes_amd checks"
    L = ctypes.CDLL(LIB)
    dp = ctypes.POINTER(ctypes.c_double)
    L.tdt_divcheck.argtypes = [dp, dp, ctypes.c_int, dp, dp, dp]
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = [np.empty_like(a) for _ in range(3)]
    rc = L.tdt_divcheck(*(x.ctypes.data_as(dp) for x in (a, b)), a.size,
                        *(x.ctypes.data_as(dp) for x in out))
    assert rc == 0, f"HIP error {rc}"
    return out


def random_doubles(rng, n, emin, emax):
    mant = 1.0 + rng.random(n)
    return np.ldexp(mant, rng.integers(emin, emax + 1, n)) * rng.choice([-1.0, 1.0], n)


def cases():
    rng = np.random.default_rng(20261203)
    a, b = [], []
    # wide random: exponents within +-280
    a.append(random_doubles(rng, 1 << 20, -280, 280)); b.append(random_doubles(rng, 1 << 20, -280, 280))
    # physical ranges: densities, momenta, energies, nu~, metric terms
    a.append(rng.uniform(-1e7, 1e7, 1 << 18)); b.append(rng.uniform(1e-3, 20.0, 1 << 18))
    a.append(rng.uniform(-1e-6, 1e-6, 1 << 18)); b.append(rng.uniform(1e-12, 1e-6, 1 << 18))
    # denominators a few ulps from 1, 3, 7, 10 and numerators near exact multiples
    for base in (1.0, 3.0, 7.0, 10.0, 0.1):
        k = rng.integers(-64, 65, 1 << 16)
        bb = base * (1.0 + k * 2.0 ** -52)
        q = np.round(rng.uniform(1.0, 2.0, 1 << 16) * 2 ** 20) / 2 ** 20
        aa = bb * q
        aa = aa + rng.integers(-3, 4, 1 << 16) * np.spacing(aa)
        a.append(aa); b.append(bb)
    # integers and small counts (nodal averages)
    ia = rng.integers(-10 ** 6, 10 ** 6, 1 << 16).astype(np.float64)
    a.append(ia); b.append(rng.integers(1, 9, 1 << 16).astype(np.float64))
    return np.concatenate(a), np.concatenate(b)


def test_recip_division_matches_ieee():
    a, b = cases()
    fast, ieee, quot = run(a, b)
    ref = a / b
    assert np.array_equal(ieee.view(np.int64), ref.view(np.int64)), "device a / b is not IEEE"
    bad = fast.view(np.int64) != ref.view(np.int64)
    assert not bad.any(), (f"{bad.sum()} of {a.size} differ, e.g. a={a[bad][:3]} b={b[bad][:3]} "
                           f"fast={fast[bad][:3]} ieee={ref[bad][:3]}")
    assert np.array_equal(quot.view(np.int64), ref.view(np.int64))


def test_recip_division_edge_values():
    """Zero numerators give zero (the sign of -0 / b, b > 0, is not kept);
    zero, infinite and NaN operands give non-finite results where a / b does."""
    a = np.array([0.0, -0.0, 0.0, -0.0, 1.0, 1.0, np.inf, np.nan, 1.0, 2.0])
    b = np.array([2.0, -2.0, -2.0, 3.0, 0.0, np.nan, 2.0, 2.0, np.inf, -np.inf])
    fast, ieee, _ = run(a, b)
    assert np.all(fast[:4] == 0.0)
    assert np.array_equal(np.signbit(fast[:3]), np.signbit(ieee[:3]))
    assert not np.isfinite(fast[4:8]).any() and not np.isfinite(ieee[4:8]).any()
    # b infinite: IEEE gives +-0, the helper NaN; a state holding inf is flagged anyway
    assert np.all(ieee[8:] == 0.0)
