"""The wider-precision oracles (long double, suffix f80; binary128, suffix
f128) and their finite differences, on the CPU.  Each is the FP64 map
evaluated with a wider significand, so it is pinned to the FP64 oracle
(itself bit-identical to the compiled reference) by closeness, and the
finite-difference routine by maps whose derivative is known:

* a 20-step rollout agrees with the FP64 oracle to a few FP64 roundings
  (relative L1), from the same state with the same dt sequence;
* a uniform moving state with periodic boundaries is a fixed point of the
  step (every face carries the same flux), so the rollout map is the
  identity along uniform directions and its finite difference returns the
  direction itself;
* the weighted form (derivative of sum(w * U_n)) equals the weighted sum of
  the field form;
* at one eps the FP64, f80 and f128 differences of a step cross the same
  branches, so they differ by their roundoff only (FP64 about 1e-16 / eps);
* with the Newton counts frozen, the f128 differences at eps 1e-8 .. 1e-12
  round to the same doubles: the derivative of the step as executed, which
  the GPU tangent and adjoint are checked against.  (Larger eps cross limiter
  and upwind kinks: on the smooth states below, eps 1e-6 is already 2e-5 off
  for first-order slopes.)
"""
import numpy as np
import pytest

from oracle_lib import IC_DIAG, U, U2, Oracle


def _smooth_state(nx, ny, seed, gamma=1.4):
    rng = np.random.default_rng(seed)
    x = (np.arange(nx + 4) - 1.5) / nx
    y = (np.arange(ny + 4) - 1.5) / ny
    X, Y = np.meshgrid(x, y, indexing="ij")

    def field():
        f = np.zeros_like(X)
        for _ in range(3):
            kx, ky = rng.integers(1, 4, 2)
            f += rng.normal() * np.sin(2 * np.pi * (kx * X + rng.uniform())) * np.cos(2 * np.pi * (ky * Y + rng.uniform()))
        return f / 3

    rho = 1.0 + 0.3 * field()
    u = 0.2 * field()
    v = 0.2 * field()
    p = 1.0 + 0.3 * field()
    return np.stack([rho, p / (gamma - 1) + 0.5 * rho * (u * u + v * v), rho * u, rho * v], axis=-1)


def _interior(nx, ny):
    m = np.zeros((nx + 4, ny + 4, 4), dtype=bool)
    m[2:nx + 2, 2:ny + 2] = True
    return m


def _rel_l1(a, b):
    return float(np.abs(a - b).sum() / np.abs(b).sum())


@pytest.mark.parametrize("wide", ["f80", "f128"])
def test_wide_map_close_to_fp64_oracle(wide):
    nx, steps = 64, 20
    o64 = Oracle(nx, precision="f64")
    o64.init(IC_DIAG)
    _, dts = o64.run(steps)
    u64 = o64.get(steps % 2)
    ow = Oracle(nx, precision=wide)
    ow.init(IC_DIAG)
    ow.compute_dt(0)
    ow.make_boundaries(0)
    for n, dt in enumerate(dts):
        ow.godunov(n, dt)
    uw = ow.get(steps % 2)
    i = slice(2, nx + 2)
    errs = [_rel_l1(uw[i, i, v], u64[i, i, v]) for v in range(4)]
    print(f"{wide} vs f64 after 20 steps of the diagonal problem, rel L1:", errs)
    assert max(errs) <= 1e-14
    assert not np.array_equal(uw, u64)  # a different precision, not the FP64 map again


def test_fd_identity_on_uniform_flow():
    nx, gamma = 24, 1.4
    o = Oracle(nx, bc=(3, 3, 3, 3), precision="f80")
    U0 = np.zeros((nx + 4, nx + 4, 4))
    rho, u, v, p = 1.3, 0.4, -0.25, 0.9
    U0[..., 0] = rho
    U0[..., 1] = p / (gamma - 1) + 0.5 * rho * (u * u + v * v)
    U0[..., 2] = rho * u
    U0[..., 3] = rho * v
    mask = _interior(nx, nx)
    # a uniform direction on every cell, ghosts and corners included (the
    # corners, never filled, are read by the flux loop as they are)
    d = np.zeros_like(U0)
    d[...] = np.array([0.3, -0.7, 0.2, 0.5])
    dts = [2e-3] * 3
    g = o.fd(U0, d, dts, 1e-6)
    err = np.abs(g[mask] - d[mask]).max()
    print("uniform flow: max |FD - direction| =", err)
    assert err <= 1e-11


def test_fd_weighted_equals_field_sum():
    nx = 32
    o = Oracle(nx, bc=(1, 3, 2, 1), precision="f80")
    U0 = _smooth_state(nx, nx, 5)
    d = np.where(_interior(nx, nx), np.random.default_rng(6).normal(size=U0.shape), 0.0)
    w = np.random.default_rng(7).normal(size=U0.shape)
    dts = [1e-3, 1.2e-3]
    field = o.fd(U0, d, dts, 1e-6)
    scalar = o.fd(U0, d, dts, 1e-6, weight=w)
    assert abs(scalar - float(np.sum(w * field))) <= 1e-12 * max(1.0, abs(scalar))


@pytest.mark.parametrize("slope_type", [0, 1, 2])
def test_fd_precisions_and_convergence(slope_type):
    """One step from a smooth state with flow (periodic)."""
    nx, dt = 48, 1.5e-3
    U0 = _smooth_state(nx, nx, 11 + slope_type)
    mask = _interior(nx, nx)
    d = np.where(mask, _smooth_state(nx, nx, 21 + slope_type) - 1.0, 0.0)
    o = {prec: Oracle(nx, bc=(3, 3, 3, 3), slope_type=slope_type, precision=prec) for prec in ("f64", "f80", "f128")}
    same_eps = {prec: o[prec].fd(U0, d, [dt], 1e-6, freeze=True)[mask] for prec in o}
    ref = same_eps["f128"]
    nrm = np.linalg.norm(ref)
    e64 = np.linalg.norm(same_eps["f64"] - ref) / nrm
    e80 = np.linalg.norm(same_eps["f80"] - ref) / nrm
    conv = {eps: o["f128"].fd(U0, d, [dt], eps, freeze=True)[mask] for eps in (1e-8, 1e-10, 1e-12)}
    spread = max(np.linalg.norm(conv[e] - conv[1e-10]) for e in conv) / np.linalg.norm(conv[1e-10])
    print(f"slope_type {slope_type}: eps 1e-6: |f64 - f128| {e64:.1e}, |f80 - f128| {e80:.1e}; "
          f"f128 at eps 1e-8..1e-12: spread {spread:.1e}")
    assert e64 <= 1e-8
    assert e80 <= 1e-11
    assert spread <= 1e-15
