"""The device division and square-root cores against IEEE arithmetic.

The FP64 and FP32 steps divide and take square roots with range-restricted
cores (reciprocal or rsqrt/sqrt estimate, Newton / Goldschmidt refinement and
remainder corrections, without the scaling and special-value fixup of the full
IEEE sequences).  Inside the solver's range (magnitudes 1e-22..1e9) they must
return the correctly rounded result: numpy's float64 / float32 division and
sqrt, bit for bit (a zero may differ in sign).  Operands: log-uniform over the
range with both signs for numerators, plus mantissa edge cases (powers of two,
all-ones mantissas, quotients next to 1, perfect squares and their neighbours),
and the hard rounding cases of tests/hard_cases.py.
"""
import numpy as np
import pytest

import euler2d_cudafortran_amd as e2d
import hard_cases as hc

pytestmark = pytest.mark.gpu

N = 1 << 21


def _edge_values(dt):
    info = np.finfo(dt)
    eps = info.eps
    base = np.array([1.0, 2.0, 0.5, 3.0, 10.0, 0.1, 1.4, 1.0 / 1.4, 7.0e-22, 1e-7, 1e9], dtype=np.float64)
    vals = [base, base * (1 + eps), base * (1 - eps / 2), base * (2 - eps), np.ldexp(1.0, np.arange(-70, 30))]
    return np.concatenate(vals).astype(dt)


def _log_uniform(rng, n, lo, hi, dt, signed=False):
    x = 10.0 ** rng.uniform(lo, hi, n)
    if signed:
        x *= rng.choice([-1.0, 1.0], n)
    return x.astype(dt)


@pytest.mark.parametrize("prec, dt", [(e2d.PREC_FP64, np.float64), (e2d.PREC_FP32, np.float32)])
def test_division_core_is_correctly_rounded(gpu_lib, prec, dt):
    rng = np.random.default_rng(11)
    a = _log_uniform(rng, N, -22, 9, dt, signed=True)
    b = _log_uniform(rng, N, -22, 9, dt, signed=True)
    e = _edge_values(dt)
    ea, eb = np.meshgrid(e, e, indexing="ij")
    a = np.concatenate([a, ea.ravel(), np.zeros(4, dt)])
    b = np.concatenate([b, eb.ravel(), np.array([1.0, 3.0, 1e-7, 1e9], dt)])
    p = e2d.HydroParams(nx=8, ny=8, precision=prec)
    got = e2d.pointwise(p, e2d.FN_DIV, np.stack([a, b], axis=1).astype(np.float64))[:, 0].astype(dt)
    want = a / b
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"{bad.size} quotients differ, first: {a[bad[:3]]} / {b[bad[:3]]} -> {got[bad[:3]]} vs {want[bad[:3]]}"


@pytest.mark.parametrize("prec, dt", [(e2d.PREC_FP64, np.float64), (e2d.PREC_FP32, np.float32)])
def test_sqrt_core_is_correctly_rounded(gpu_lib, prec, dt):
    rng = np.random.default_rng(12)
    x = _log_uniform(rng, N, -22, 12, dt)
    e = _edge_values(dt)
    k = np.arange(1, 5000, dtype=np.float64)
    sq = (k * k).astype(dt)
    eps = np.finfo(dt).eps
    x = np.concatenate([x, e, sq, (sq * (1 + eps)).astype(dt), (sq * (1 - eps / 2)).astype(dt), np.zeros(1, dt)])
    p = e2d.HydroParams(nx=8, ny=8, precision=prec)
    got = e2d.pointwise(p, e2d.FN_SQRT, x.astype(np.float64))[:, 0].astype(dt)
    want = np.sqrt(x)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"{bad.size} roots differ, first: {x[bad[:3]]} -> {got[bad[:3]]} vs {want[bad[:3]]}"


# Hard rounding cases (tests/hard_cases.py): exact quotients and roots within a
# tiny fraction of an ulp of a midpoint, spread over the solver's magnitudes.
HARD_DIV = 1 << 17
HARD_SQRT = 1 << 15


@pytest.mark.parametrize("prec, dt, p", [(e2d.PREC_FP64, np.float64, 53), (e2d.PREC_FP32, np.float32, 24)])
def test_division_core_hard_cases(gpu_lib, prec, dt, p):
    a, b = hc.scaled_div_operands(hc.div_near_midpoint(HARD_DIV, p, seed=21), p, seed=22)
    p_ = e2d.HydroParams(nx=8, ny=8, precision=prec)
    got = e2d.pointwise(p_, e2d.FN_DIV, np.stack([a, b], axis=1))[:, 0].astype(dt)
    want = a.astype(dt) / b.astype(dt)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"{bad.size} of {a.size} hard quotients differ, first: {a[bad[:3]]} / {b[bad[:3]]}"


@pytest.mark.parametrize("prec, dt, p", [(e2d.PREC_FP64, np.float64, 53), (e2d.PREC_FP32, np.float32, 24)])
def test_sqrt_core_hard_cases(gpu_lib, prec, dt, p):
    x = hc.scaled_radicands(hc.sqrt_near_midpoint(HARD_SQRT, p, seed=23), p, seed=24)
    p_ = e2d.HydroParams(nx=8, ny=8, precision=prec)
    got = e2d.pointwise(p_, e2d.FN_SQRT, x)[:, 0].astype(dt)
    want = np.sqrt(x.astype(dt))
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"{bad.size} of {x.size} hard roots differ, first: {x[bad[:3]]}"


@pytest.mark.parametrize("prec, dt, p", [(e2d.PREC_FP64, np.float64, 53), (e2d.PREC_FP32, np.float32, 24)])
def test_division_by_root_is_correctly_rounded(gpu_lib, prec, dt, p):
    """a / sqrt(x) as the two-shock solver forms (p* - p) / w (E2D_FN_DIV_ROOT):
    in FP64 the quotient takes its reciprocal from the square root's own
    Goldschmidt iteration instead of a fresh one; root and quotient must still be
    the IEEE results, a / RN(sqrt(x)) rounded once."""
    rng = np.random.default_rng(25)
    a = _log_uniform(rng, N, -22, 9, np.float64, signed=True)
    x = _log_uniform(rng, N, -22, 12, np.float64)
    hx = np.concatenate([_log_uniform(rng, HARD_DIV, -22, 12, np.float64),
                         hc.scaled_radicands(hc.sqrt_near_midpoint(HARD_SQRT, p, seed=26), p, seed=27)])
    ha, hx = hc.div_by_root_operands(hx, p, seed=28)
    a = np.concatenate([a, ha]).astype(dt)
    x = np.concatenate([x, hx]).astype(dt)
    p_ = e2d.HydroParams(nx=8, ny=8, precision=prec)
    y = e2d.pointwise(p_, e2d.FN_DIV_ROOT, np.stack([a, x], axis=1).astype(np.float64)).astype(dt)
    s = np.sqrt(x)
    bad_s = np.nonzero(y[:, 1] != s)[0]
    assert bad_s.size == 0, f"{bad_s.size} roots differ, first: {x[bad_s[:3]]}"
    want = a / s
    bad = np.nonzero(y[:, 0] != want)[0]
    assert bad.size == 0, (f"{bad.size} of {a.size} quotients differ ({ha.size} hard), first: "
                           f"{a[bad[:3]]} / sqrt({x[bad[:3]]})")


def _ulps(got, want):
    """|got - want| in units of the last place of want (float32)."""
    spacing = np.abs(np.spacing(want.astype(np.float32))).astype(np.float64)
    return np.abs(got.astype(np.float64) - want.astype(np.float64)) / spacing


def test_native_float_arithmetic_error(gpu_lib):
    """fp32_native's native reciprocal and square root: v_rcp_f32 then one
    multiplication for a / b, v_sqrt_f32 for sqrt, each within about an ulp."""
    rng = np.random.default_rng(29)
    a = _log_uniform(rng, N, -22, 9, np.float32, signed=True)
    b = _log_uniform(rng, N, -22, 9, np.float32, signed=True)
    x = _log_uniform(rng, N, -22, 12, np.float32)
    p_ = e2d.HydroParams(nx=8, ny=8, precision=e2d.PREC_FP32_NATIVE)
    q = e2d.pointwise(p_, e2d.FN_DIV, np.stack([a, b], axis=1).astype(np.float64))[:, 0]
    r = e2d.pointwise(p_, e2d.FN_SQRT, x.astype(np.float64))[:, 0]
    dq = _ulps(q, a / b)
    dr = _ulps(r, np.sqrt(x))
    print(f"fp32_native: division max {dq.max():.3f} ulp ({np.mean(dq > 0):.3f} inexact), "
          f"sqrt max {dr.max():.3f} ulp ({np.mean(dr > 0):.3f} inexact)")
    assert dq.max() <= 2.0 and dr.max() <= 1.0


def test_division_from_a_near_reciprocal_is_correctly_rounded(gpu_lib):
    """a / b with the reciprocal of b refined by one Newton step from a start r0
    within about 2^-50 of 1/b (E2D_FN_DIV_NEAR), as the FP64 two-shock finish
    forms its quotients by wo^2 (start (1/wo)^2) and by r* (start (1 + E)/ro):
    the IEEE quotient for every start offset up to 8 ulp, on random and hard
    operands."""
    rng = np.random.default_rng(31)
    a = _log_uniform(rng, N, -22, 9, np.float64, signed=True)
    b = _log_uniform(rng, N, -22, 9, np.float64, signed=True)
    ha, hb = hc.scaled_div_operands(hc.div_near_midpoint(HARD_DIV, 53, seed=32), 53, seed=33)
    a, b = np.concatenate([a, ha]), np.concatenate([b, hb])
    k = rng.integers(-8, 9, a.size).astype(np.float64)
    r0 = (1.0 / b) * (1.0 + k * 2.0 ** -53)
    p_ = e2d.HydroParams(nx=8, ny=8, precision=e2d.PREC_FP64)
    got = e2d.pointwise(p_, e2d.FN_DIV_NEAR, np.stack([a, b, r0], axis=1))[:, 0]
    want = a / b
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"{bad.size} of {a.size} quotients differ, first: {a[bad[:3]]} / {b[bad[:3]]} from {r0[bad[:3]]}"


@pytest.mark.parametrize("prec, dt", [(e2d.PREC_FP64, np.float64), (e2d.PREC_FP32, np.float32)])
def test_newton_stop_test_is_exact(gpu_lib, prec, dt):
    """The two-shock solver's stop test, |delp| / (p* + smallpp) > conv_tol
    (E2D_FN_CONV_GT).  FP64 and FP32 decide it with one fma (in double) and a
    comparison against the rounding boundary above conv_tol instead of forming
    the quotient; the decision must be that of the IEEE quotient in the
    precision, on quotients a tiny fraction of an ulp either side of that
    boundary and on random ones around conv_tol."""
    tol = float(np.float32(1e-6))
    rng = np.random.default_rng(35)
    d = _log_uniform(rng, N, -8, 6, np.float64)
    a = d * tol * (1.0 + rng.uniform(-1e-6, 1e-6, N))
    a[: N // 8] = 0.0
    a, d = a.astype(dt).astype(np.float64), d.astype(dt).astype(np.float64)
    want = (a.astype(dt) / d.astype(dt)) > dt(tol)
    p = 53 if dt is np.float64 else 24
    ha, hd, above = hc.quotients_next_to_midpoint(tol, HARD_DIV, seed=36, max_r=1 << (20 if p == 53 else 6), p=p)
    assert np.array_equal((ha.astype(dt) / hd.astype(dt)) > dt(tol), above)
    a, d, want = np.concatenate([a, ha]), np.concatenate([d, hd]), np.concatenate([want, above])
    p_ = e2d.HydroParams(nx=8, ny=8, precision=prec)
    got = e2d.pointwise(p_, e2d.FN_CONV_GT, np.stack([a, d], axis=1))[:, 0] == 1.0
    bad = np.nonzero(got != want)[0]
    assert 0.2 < want.mean() < 0.6
    assert bad.size == 0, f"{bad.size} of {a.size} decisions differ, first: {a[bad[:3]]} / {d[bad[:3]]}"
