"""The device division and square-root cores against IEEE arithmetic.

The FP64 and FP32 steps divide and take square roots with range-restricted
cores (reciprocal or rsqrt/sqrt estimate, Newton / Goldschmidt refinement and
remainder corrections, without the scaling and special-value fixup of the full
IEEE sequences).  Inside the solver's range (magnitudes 1e-22..1e9) they must
return the correctly rounded result: numpy's float64 / float32 division and
sqrt, bit for bit (a zero may differ in sign).  Operands: log-uniform over the
range with both signs for numerators, plus mantissa edge cases (powers of two,
all-ones mantissas, quotients next to 1, perfect squares and their neighbours).
"""
import numpy as np
import pytest

import euler2d_cudafortran_amd as e2d

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
