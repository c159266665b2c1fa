"""The hard rounding cases of tests/hard_cases.py, checked in exact arithmetic.

Each construction must put the exact quotient or root where its docstring
says (next to a midpoint, on the stated side), and numpy's IEEE division and
square root must round those cases the way the exact value says; the GPU
tests then take numpy's results as the expected ones.
"""
from fractions import Fraction

import numpy as np
import pytest

import hard_cases as hc


@pytest.mark.parametrize("p, dt", [(53, np.float64), (24, np.float32)])
def test_division_cases_sit_next_to_midpoints(p, dt):
    cases = hc.div_near_midpoint(2000, p, seed=1)
    for a, b, m, r in cases:
        exact = Fraction(a, b)
        assert exact - Fraction(m, 1 << p) == Fraction(r, (1 << p) * b)
        assert m & 1 and (1 << p) <= m < (1 << (p + 1))
        want = Fraction(m + (1 if r > 0 else -1), 1 << p)  # the neighbour on r's side
        got = dt(a) / dt(b)
        assert Fraction(float(got)) == want
    # scaled copies round alike
    sa, sb = hc.scaled_div_operands(cases, p, seed=2)
    q = sa.astype(dt) / sb.astype(dt)
    base = np.array([c[0] for c in cases], dt) / np.array([c[1] for c in cases], dt)
    mq, _ = np.frexp(np.abs(q))
    mb, _ = np.frexp(base)
    assert np.array_equal(mq, mb)


@pytest.mark.parametrize("p, dt", [(53, np.float64), (24, np.float32)])
def test_sqrt_cases_sit_next_to_midpoints(p, dt):
    cases = hc.sqrt_near_midpoint(1000, p, seed=3)
    assert len(cases) == 1000
    for big, b, m, t in cases:
        x = Fraction(big, 1 << (p - 1 - b))
        assert 1 <= x < 4 and (1 << (p - 1)) <= big < (1 << p)
        mid = Fraction(m, 1 << p)
        assert mid * mid - x == Fraction(t, 1 << (2 * p))
        # the root is within |t| / 2^(p+1) of the midpoint, in ulps of 2^(1-p)
        want = Fraction(m + (-1 if t > 0 else 1), 1 << p)
        got = np.sqrt(dt(float(x)))
        assert Fraction(float(got)) == want
    xs = hc.scaled_radicands(cases, p, seed=4)
    base = np.array([float(Fraction(c[0], 1 << (p - 1 - c[1]))) for c in cases])
    m1, _ = np.frexp(np.sqrt(xs.astype(dt)))
    m2, _ = np.frexp(np.sqrt(base.astype(dt)))
    assert np.array_equal(m1, m2)


@pytest.mark.parametrize("p, dt", [(53, np.float64), (24, np.float32)])
def test_division_by_root_cases_sit_next_to_midpoints(p, dt):
    rng = np.random.default_rng(5)
    x = 10.0 ** rng.uniform(-22, 9, 3000)
    a, xr = hc.div_by_root_operands(x, p, seed=6)
    assert len(a) > 500
    assert np.array_equal(xr.astype(dt).astype(np.float64), xr)
    s = np.sqrt(xr.astype(dt))
    for av, sv in zip(a[:300].tolist(), s[:300].tolist()):
        exact = abs(Fraction(av) / Fraction(float(sv)))
        # nearest midpoint of the quotient's binade, and the distance to it
        e = exact.numerator.bit_length() - exact.denominator.bit_length()
        if Fraction(2) ** e > exact:
            e -= 1
        unit = Fraction(2) ** (e - p)          # half an ulp of the quotient
        k = exact / unit
        mid = (int(k) | 1) if int(k) % 2 == 1 else int(k) + 1
        nearest_mid = min((mid - 2, mid, mid + 2), key=lambda j: abs(k - j))
        assert abs(k - nearest_mid) < Fraction(1, 1 << (p - 6))


@pytest.mark.parametrize("p, dt, bound", [(53, np.float64, 70), (24, np.float32, 36)])
def test_stop_test_cases_sit_next_to_the_tolerance_midpoint(p, dt, bound):
    tol = dt(np.float32(1e-6))
    a, d, above = hc.quotients_next_to_midpoint(float(tol), 2000, seed=7, max_r=1 << (20 if p == 53 else 6), p=p)
    assert np.array_equal(a.astype(dt).astype(np.float64), a) and np.array_equal(d.astype(dt).astype(np.float64), d)
    h = np.spacing(tol) / 2
    mu = Fraction(float(tol)) + Fraction(float(h))
    assert Fraction(float(np.nextafter(tol, dt(1.0)))) - Fraction(float(tol)) == 2 * Fraction(float(h))
    for av, dv, up in zip(a.tolist(), d.tolist(), above.tolist()):
        q = Fraction(av) / Fraction(dv)
        # within 2^-bound of mu relative: a small fraction of an ulp from the rounding boundary
        assert (q > mu) == up and abs(q - mu) < mu * Fraction(1, 1 << bound)
        assert ((dt(av) / dt(dv)) > tol) == up
    assert 0.3 < above.mean() < 0.7
