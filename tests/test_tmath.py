"""Accuracy of the hot path's FP64 exp / log / pow / tanh (csrc/tmath.hpp)
against glibc, on the CPU: the same header is compiled for the host."""
import os
import subprocess

import pytest

from conftest import PKG, REPO

BOUNDS = {"tanh_small": 4, "tanh_mid": 4, "tanh_neg": 4, "pow_0.6": 3, "pow_-2.5": 5,
          "pow_1/7": 3, "pow_3.5": 6, "exp": 2, "log": 2}


def test_tmath_within_a_few_ulps_of_glibc(tmp_path):
    exe = str(tmp_path / "tmath_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                           "-I" + os.path.join(PKG, "csrc"),
                           os.path.join(REPO, "tests", "native", "tmath_check.cpp"), "-o", exe,
                           "-lm"])
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    rows = {ln.split()[0]: ln.split()[1:] for ln in out.strip().splitlines()}
    assert int(rows["special"][0]) == 0, "special values (0, inf, NaN, negative base) wrong"
    for name, bound in BOUNDS.items():
        mx, mean = float(rows[name][0]), float(rows[name][1])
        assert mx <= bound, f"{name}: max error {mx} ulp > {bound}"
        assert mean <= 1.0, f"{name}: mean error {mean} ulp"
