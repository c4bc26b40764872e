"""CPU: the Boys function pieces.

* the oracle's restatement of the reference's three-branch scheme
  (generator_boys_function.py:98-205) reproduces the reference's own tables at
  grid points, matches mpmath where the scheme is accurate, and matches the
  closed-form asymptotic series in its large-x branch (where the reference
  drops the exp(-x) term);
* the device's grid tables (intception_amd/boys_tables.py, mpmath) agree with
  the reference's tables; the device Chebyshev table is accurate to 1e-15.
"""
import math
import os

import mpmath as mp
import numpy as np
import pytest

from intception_amd import boys_tables
from oracle import oracle

TAB = np.load(os.path.join(os.path.dirname(__file__), "golden", "boys_reference_tables.npz"))


def boys_mp(n, x):
    return float(boys_tables.boys_mp(n, x))


def test_ref_scheme_reproduces_table_at_grid_points():
    for k in (0, 7, 55, 150):
        x = k / 10.0
        f = oracle.boys_ref(20, x)
        assert abs(f[20] - TAB["smallx"][k][20]) <= 2e-16 * TAB["smallx"][k][20]  # dx ~ 0: the table entry
        for n in range(20):
            assert abs(f[n] - boys_mp(n, x)) <= 2e-14 * boys_mp(n, x)


def test_ref_scheme_accuracy_where_accurate():
    for x in np.linspace(0.0, 34.99, 157):
        f = oracle.boys_ref(8, float(x))
        for n in range(9):
            r = boys_mp(n, x)
            assert abs(f[n] - r) <= 2e-13 * r, (x, n)
    for x in (60.0, 75.5, 120.0):
        f = oracle.boys_ref(8, x)
        for n in range(9):
            r = boys_mp(n, x)
            assert abs(f[n] - r) <= 1e-13 * r, (x, n)


def test_ref_scheme_large_x_branch_is_asymptotic_series():
    # F_n(x) ~ (2n-1)!! / 2^(n+1) sqrt(pi / x^(2n+1)) exactly as emitted
    for x in (35.0, 41.3, 57.0, 99.0):
        f = oracle.boys_ref(12, x)
        for n in range(13):
            df = 1
            for k in range(2 * n - 1, 0, -2):
                df *= k
            asym = df / 2 ** (n + 1) * math.sqrt(math.pi / x ** (2 * n + 1))
            assert abs(f[n] - asym) <= 1e-14 * asym


def test_ref_scheme_known_large_x_defect():
    # documented inaccuracy of the reference for 35 <= x < 60 and high n
    # (measured: x = 36 -> 9.7e-9 relative at n = 8, 1.9e-6 at n = 12)
    f = oracle.boys_ref(12, 36.0)
    assert abs(f[8] - boys_mp(8, 36.0)) / boys_mp(8, 36.0) > 1e-9
    assert abs(f[12] - boys_mp(12, 36.0)) / boys_mp(12, 36.0) > 1e-7


def test_exact_mode():
    for x in list(np.linspace(0, 50, 101)) + [0.013, 29.99, 30.0, 30.01, 80.0]:
        f = oracle.boys_exact(16, float(x))
        for n in range(17):
            r = boys_mp(n, x)
            assert abs(f[n] - r) <= 2e-14 * r, (x, n)


def test_device_grid_tables_match_reference_tables():
    small, medium = boys_tables.grid_tables()
    small, medium = np.array(small), np.array(medium)
    assert small.shape == TAB["smallx"].shape and medium.shape == TAB["mediumx"].shape
    assert np.max(np.abs(small - TAB["smallx"]) / np.abs(TAB["smallx"])) < 5e-16
    assert np.max(np.abs(medium - TAB["mediumx"]) / np.abs(TAB["mediumx"])) < 1e-13


@pytest.mark.parametrize("m", [0, 3, 8, 12, 16])
def test_device_chebyshev_table(m):
    mp.mp.dps = 30
    worst = 0.0
    for k in range(0, boys_tables.NINT, 5):
        poly = [float(v) for v in boys_tables.cheb_interval_poly(m, k)]
        for t in np.linspace(-1, 1, 7):
            v = 0.0
            for c in reversed(poly):
                v = v * t + c
            x = k + (t + 1) / 2
            r = boys_mp(m, x)
            worst = max(worst, abs(v - r) / r)
    assert worst < 2e-15
