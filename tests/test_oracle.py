"""The oracle (CPU restatement, oracle/cpu_ref*.c) pinned against the reference.

* every golden fixture (outputs of the reference compiled from serial_ref/,
  tests/golden/make_golden.py) is reproduced bit for bit;
* closed forms the mathematics fixes: the flux of a uniform state, the exact
  star pressure of a two-shock problem (Toro), textbook minmod / MC limiters,
  the first-order (slope_type 0) step assembled from pointwise fluxes, and the
  exact transpose symmetry of the scheme on a symmetric initial condition.
CPU only.
"""
import hashlib

import numpy as np
import pytest

from oracle_lib import (FX, FY, IC_DIAG, IC_IMPLODE, KAT_CMPFLX, KAT_ITERS, KAT_PRIM, KAT_RIEMANN,
                        KAT_SLOPE, Q, U, U2, Oracle)

KINDS = ("fp64", "fp32")


def _same(a, b):
    """Bitwise agreement up to the sign of zero (== semantics); a NaN must
    sit where the reference has one (its HLLC solver divides by zero on some
    degenerate states)."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return False
    na, nb = np.isnan(a), np.isnan(b)
    return bool(np.array_equal(na, nb) and np.all(a[~na] == b[~nb]))


def test_manifest(golden):
    for line in (golden / "MANIFEST.sha256").read_text().split("\n"):
        if not line.strip():
            continue
        digest, name = line.split()
        assert hashlib.sha256((golden / name).read_bytes()).hexdigest() == digest, name


def _kat_files(golden):
    return sorted(golden.glob("kat_*.npz"))


@pytest.mark.parametrize("kind", KINDS)
def test_kat_fixtures(golden, kind):
    files = [f for f in _kat_files(golden) if f.stem.endswith(kind)]
    assert len(files) == 11
    for f in files:
        d = np.load(f)
        o = Oracle(8, slope_type=int(d["slope_type"]), niter_riemann=int(d["niter"]),
                   gamma0=float(d["gamma0"]), smallr=float(d["smallr"]), smallc=float(d["smallc"]),
                   precision="f64" if kind == "fp64" else "f32")
        y = o.kat(int(d["func"]), d["x"])
        assert _same(y, d["y"]), f.name


def test_kat_coverage(golden):
    """The Riemann fixtures exercise every branch that decides the result."""
    d = np.load(golden / "kat_riemann_fp64.npz")
    o = Oracle(8)
    iters = o.kat(KAT_ITERS, d["x"])[:, 0]
    assert iters.min() == 1 and iters.max() >= 5
    x, y = d["x"], d["y"]
    rl, rr = x[:, 0], x[:, 4]
    assert (rl < 1e-7).any() and (rr < 1e-7).any()      # density floor
    assert (y[:, 2] > 0).any() and (y[:, 2] < 0).any()  # both contact directions
    d1 = np.load(golden / "kat_riemann_niter1_fp64.npz")
    assert np.load(golden / "kat_riemann_floor_fp64.npz")["x"].shape[0] == 1000
    assert int(d1["niter"]) == 1


def _run_files(golden):
    return sorted(golden.glob("run_*.npz"))


@pytest.mark.parametrize("name", [p.stem[4:] for p in sorted(
    __import__("pathlib").Path(__file__).parent.joinpath("golden").glob("run_*.npz"))])
def test_run_fixtures(golden, name):
    d = np.load(golden / f"run_{name}.npz")
    nx, ny, b0, b1, b2, b3, st, steps, ic = (int(v) for v in d["params"])
    prec = "f32" if name.endswith("fp32") or "_fp32_" in name else "f64"
    o = Oracle(nx, ny, bc=(b0, b1, b2, b3), slope_type=st, gamma0=float(d["gamma0"]), precision=prec)
    o.init(IC_DIAG if ic == 1 else IC_IMPLODE)
    n, dts = o.run(steps)
    assert n == steps
    assert _same(dts, d["dt"])
    assert _same(o.get(U), d["u_end"])
    assert _same(o.get(U2), d["u2_end"])


def test_run_fixture_snapshots_and_stages(golden):
    d = np.load(golden / "run_diag128_fp64.npz")
    o = Oracle(128)
    o.init(IC_DIAG)
    dt = o.compute_dt(0)
    o.make_boundaries(0)
    for n in range(10):
        dt = o.compute_dt(n % 2)
        assert dt == d["dt"][n]
        o.godunov(n, dt)
        if n == 0:
            assert _same(o.get(Q), d["stage0_q"])
            assert _same(o.get(FX), d["stage0_fx"])
            assert _same(o.get(FY), d["stage0_fy"])
        if f"snap_{n}" in d:
            assert _same(o.get(U2 if n % 2 == 0 else U), d[f"snap_{n}"])


# ---------------------------------------------------------------------------
# closed forms

def _euler_flux(q, gamma):
    r, p, u, v = q
    E = p / (gamma - 1) + 0.5 * r * (u * u + v * v)
    return np.array([r * u, u * (E + p), r * u * u + p, r * u * v])


def test_uniform_state_flux_and_slopes():
    rng = np.random.default_rng(1)
    o = Oracle(8)
    q = np.column_stack([rng.uniform(0.1, 5, 200), rng.uniform(0.1, 5, 200),
                         rng.uniform(-2, 2, 200), rng.uniform(-2, 2, 200)])
    y = o.kat(KAT_RIEMANN, np.concatenate([q, q], axis=1))
    exact = np.array([_euler_flux(qq, 1.4) for qq in q])
    assert np.allclose(y[:, 4:], exact, rtol=1e-12, atol=1e-13)
    assert np.allclose(y[:, :4], q, rtol=1e-12, atol=1e-13)
    # cmpflx alone is the Euler flux of the state
    assert np.allclose(o.kat(KAT_CMPFLX, q), exact, rtol=1e-14, atol=1e-14)
    # slopes of a uniform neighbourhood vanish
    s = o.kat(KAT_SLOPE, np.concatenate([q] * 5, axis=1))
    assert np.all(s == 0)


def test_prim_closed_form():
    rng = np.random.default_rng(2)
    o = Oracle(8)
    r = rng.uniform(0.1, 5, 300)
    u, v, p = rng.uniform(-2, 2, 300), rng.uniform(-2, 2, 300), rng.uniform(0.1, 5, 300)
    E = p / 0.4 + 0.5 * r * (u * u + v * v)
    y = o.kat(KAT_PRIM, np.column_stack([r, E, r * u, r * v]))
    assert np.allclose(y[:, 0], r) and np.allclose(y[:, 2], u, rtol=1e-14)
    assert np.allclose(y[:, 1], p, rtol=1e-12)
    assert np.allclose(y[:, 4], np.sqrt(1.4 * p / r), rtol=1e-12)


def _toro_pstar(rl, ul, pl, rr, ur, pr, g):
    """Exact star pressure (Toro, Riemann Solvers..., ch. 4), Newton on f(p)."""
    def fk(p, r, pk):
        a = 2 / ((g + 1) * r)
        b = (g - 1) / (g + 1) * pk
        ck = np.sqrt(g * pk / r)
        if p > pk:
            f = (p - pk) * np.sqrt(a / (p + b))
            df = np.sqrt(a / (b + p)) * (1 - (p - pk) / (2 * (b + p)))
        else:
            f = 2 * ck / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1)
            df = 1 / (r * ck) * (p / pk) ** (-(g + 1) / (2 * g))
        return f, df
    p = 0.5 * (pl + pr)
    for _ in range(100):
        f1, d1 = fk(p, rl, pl)
        f2, d2 = fk(p, rr, pr)
        dp = (f1 + f2 + ur - ul) / (d1 + d2)
        p = max(p - dp, 1e-12)
        if abs(dp) < 1e-15 * p:
            break
    return p


def test_two_shock_star_state_matches_exact():
    """For colliding flows both waves are shocks: the two-shock Newton
    iteration solves the exact Riemann problem.  Where both shocks move away
    from x/t = 0 the Godunov state is the star state, so p*, u* and the
    Rankine-Hugoniot star density of the upwind side agree with Toro's."""
    g = 1.4
    rng = np.random.default_rng(3)
    o = Oracle(8)
    checked = 0
    for _ in range(400):
        rl, rr = rng.uniform(0.2, 3, 2)
        pl, pr = rng.uniform(0.2, 3, 2)
        ul = rng.uniform(0.2, 2)
        ur = -rng.uniform(0.2, 2)
        ps = _toro_pstar(rl, ul, pl, rr, ur, pr, g)
        if ps <= max(pl, pr) * 1.01:
            continue
        al, ar = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
        sl = ul - al * np.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
        sr = ur + ar * np.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
        if not (sl < -0.05 and sr > 0.05):
            continue
        fl = (ps - pl) * np.sqrt(2 / ((g + 1) * rl) / (ps + (g - 1) / (g + 1) * pl))
        fr = (ps - pr) * np.sqrt(2 / ((g + 1) * rr) / (ps + (g - 1) / (g + 1) * pr))
        us = 0.5 * (ul + ur) + 0.5 * (fr - fl)
        rk, pk = (rl, pl) if us >= 0 else (rr, pr)
        k = (g - 1) / (g + 1)
        rs = rk * (ps / pk + k) / (k * ps / pk + 1)
        y = o.kat(KAT_RIEMANN, np.array([[rl, pl, ul, 0.3, rr, pr, ur, -0.2]]))[0]
        assert abs(y[1] - ps) <= 1e-9 * ps
        assert abs(y[2] - us) <= 1e-8 * max(1.0, abs(us))
        assert abs(y[0] - rs) <= 1e-8 * rs
        checked += 1
    assert checked > 50


def _minmod(a, b):
    return np.where(a * b <= 0, 0.0, np.sign(a) * np.minimum(np.abs(a), np.abs(b)))


def test_limiters_textbook():
    rng = np.random.default_rng(4)
    n = 500
    qs = rng.normal(0, 1, (n, 3))
    for st in (1, 2):
        o = Oracle(8, slope_type=st)
        x = np.zeros((n, 20))
        x[:, 0] = qs[:, 1]           # q
        x[:, 4] = qs[:, 2]           # q+x
        x[:, 8] = qs[:, 0]           # q-x
        x[:, 12] = x[:, 16] = qs[:, 1]
        dq = o.kat(KAT_SLOPE, x)[:, 0]
        dl, dr, dc = qs[:, 1] - qs[:, 0], qs[:, 2] - qs[:, 1], 0.5 * (qs[:, 2] - qs[:, 0])
        if st == 1:
            ref = _minmod(dl, dr)  # |dc| never below min(|dl|,|dr|) when signs agree
        else:
            ref = _minmod(_minmod(2 * dl, 2 * dr), dc)  # monotonized central
        assert np.allclose(dq, ref, rtol=0, atol=1e-15)


def test_first_order_step_from_pointwise_fluxes():
    """slope_type 0: face states are the cell states, so one step is
    u' = u + dt/dx (F(q_{i-1},q_i) - F(q_i,q_{i+1})) + y-terms, assembled here
    from pointwise Riemann calls and compared with the oracle's step."""
    nx, ny = 12, 10
    o = Oracle(nx, ny, slope_type=0, bc=(1, 2, 3, 1))
    o.init(IC_DIAG)
    u0 = o.get(U).copy()
    rng = np.random.default_rng(5)
    u0[:, :, 0] *= rng.uniform(0.8, 1.2, u0.shape[:2])
    u0[:, :, 2] = rng.uniform(-0.3, 0.3, u0.shape[:2]) * u0[:, :, 0]
    u0[:, :, 3] = rng.uniform(-0.3, 0.3, u0.shape[:2]) * u0[:, :, 0]
    u0[:, :, 1] += 0.5 * (u0[:, :, 2] ** 2 + u0[:, :, 3] ** 2) / u0[:, :, 0]
    o.set(U, u0)
    dt = o.compute_dt(0)
    o.godunov(0, dt)
    q = o.get(Q)
    uin = o.get(U)  # boundary-filled input
    out = o.get(U2)
    dtdx, dtdy = dt / (1.0 / nx), dt / (1.0 / ny)  # dt/dx with dx = 1/nx, as the reference rounds it
    S, J = nx + 4, ny + 4
    fx = np.zeros((S, J, 4))
    fy = np.zeros((S, J, 4))
    pairs_x, pairs_y, idx = [], [], []
    for j in range(2, J - 1):
        for i in range(2, S - 1):
            idx.append((i, j))
            pairs_x.append(np.concatenate([q[i - 1, j], q[i, j]]))
            l, r = q[i, j - 1][[0, 1, 3, 2]], q[i, j][[0, 1, 3, 2]]
            pairs_y.append(np.concatenate([l, r]))
    fxs = o.kat(KAT_RIEMANN, np.array(pairs_x))[:, 4:]
    fys = o.kat(KAT_RIEMANN, np.array(pairs_y))[:, 4:][:, [0, 1, 3, 2]]
    for k, (i, j) in enumerate(idx):
        fx[i, j] = fxs[k] * dtdx
        fy[i, j] = fys[k] * dtdy
    assert _same(fx, o.get(FX)) and _same(fy, o.get(FY))
    ref = out.copy()
    for j in range(2, ny + 2):
        for i in range(2, nx + 2):
            ref[i, j] = uin[i, j] + ((fy[i, j] - fy[i, j + 1]) + (fx[i, j] - fx[i + 1, j]))
    assert _same(ref, out)


def test_transpose_symmetry():
    """The diagonal initial condition is symmetric under x<->y; the scheme
    handles y by swapping the velocity components, so the result stays
    exactly symmetric: rho = rho^T, E = E^T, rho*u = (rho*v)^T."""
    o = Oracle(48)
    o.init(IC_DIAG)
    o.run(60)
    u = o.get(U)
    assert _same(u[:, :, 0], u[:, :, 0].T)
    assert _same(u[:, :, 1], u[:, :, 1].T)
    assert _same(u[:, :, 2], u[:, :, 3].T)
