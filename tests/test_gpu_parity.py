"""Parity of the gfx950 path with the reference, through the C-ABI.

Bar: FP64 and FP32 (the reference's two kinds) are bit-identical to the
reference (golden fixtures from the compiled serial_ref, and the oracle on
other sizes), up to the sign of zero; the mixed variant is checked against
FP64 with a stated tolerance.  All three implementation versions, both loops (the
host-driven compute_dt + godunov_unsplit of serial_ref/main.f90:53-76, and the
device-resident e2d_run_steps) are covered.
"""
import numpy as np
import pytest

import euler2d_cudafortran_amd as e2d
from oracle_lib import FX, FY, IC_DIAG, IC_IMPLODE, Q, U, U2, Oracle, rel_l1

pytestmark = pytest.mark.gpu


def _same(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return False
    na, nb = np.isnan(a), np.isnan(b)
    return bool(np.array_equal(na, nb) and np.all(a[~na] == b[~nb]))


def _diff_report(a, b):
    d = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b))))
    return f"{len(d)} differing entries, first {d[:5].tolist()}"


def _params(nx, ny=None, impl=0, prec=e2d.PREC_FP64, **kw):
    p = e2d.HydroParams(nx=nx, ny=nx if ny is None else ny, implementation_version=impl, precision=prec)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


# ---------------------------------------------------------------------------
# pointwise device numerics vs the reference's routines

@pytest.mark.parametrize("kind", ["fp64", "fp32"])
def test_pointwise_kat(gpu_lib, golden, kind):
    files = sorted(golden.glob(f"kat_*_{kind}.npz"))
    assert len(files) == 11
    prec = e2d.PREC_FP64 if kind == "fp64" else e2d.PREC_FP32
    for f in files:
        d = np.load(f)
        p = _params(8, prec=prec, slope_type=int(d["slope_type"]), niter_riemann=int(d["niter"]),
                    gamma0=float(d["gamma0"]), smallr=float(d["smallr"]), smallc=float(d["smallc"]))
        y = e2d.pointwise(p, int(d["func"]), d["x"])
        assert _same(y, d["y"]), f"{f.name}: {_diff_report(y, d['y'])}"


# ---------------------------------------------------------------------------
# whole runs vs the reference fixtures

def _fixture_case(golden, name):
    d = np.load(golden / f"run_{name}.npz")
    nx, ny, b0, b1, b2, b3, st, steps, ic = (int(v) for v in d["params"])
    prec = e2d.PREC_FP32 if "fp32" in name else e2d.PREC_FP64
    kw = dict(boundary_type_xmin=b0, boundary_type_xmax=b1, boundary_type_ymin=b2,
              boundary_type_ymax=b3, slope_type=st, gamma0=float(d["gamma0"]))
    return d, nx, ny, steps, ("diag" if ic == 1 else "implode"), prec, kw


RUNS = ["diag128_fp64", "implode128_fp64", "diag128_fp32", "implode64_fp32_mc",
        "diag48x40_fp64_absorb_mc", "implode40x48_fp64_periodic", "diag32_fp64_mixedbc_st0",
        "implode50x30_fp64_mixedbc"]


@pytest.mark.parametrize("impl", [0, 1, 2])
@pytest.mark.parametrize("name", RUNS)
def test_run_fixture_host_loop(gpu_lib, golden, name, impl):
    d, nx, ny, steps, ic, prec, kw = _fixture_case(golden, name)
    with e2d.HydroRun(_params(nx, ny, impl, prec, **kw)) as run:
        run.init(ic)
        dts = run.run_reference_loop(steps)
        assert _same(np.array(dts), d["dt"]), _diff_report(np.array(dts), d["dt"])
        u, u2 = run.download(0), run.download(1)
    assert _same(u, d["u_end"]), _diff_report(u, d["u_end"])
    assert _same(u2, d["u2_end"]), _diff_report(u2, d["u2_end"])


@pytest.mark.parametrize("impl", [0, 1, 2])
@pytest.mark.parametrize("name", ["diag128_fp64", "diag128_fp32", "implode40x48_fp64_periodic"])
def test_run_fixture_device_loop(gpu_lib, golden, name, impl):
    """e2d_run_steps keeps dt on the device; results equal the host loop's."""
    d, nx, ny, steps, ic, prec, kw = _fixture_case(golden, name)
    with e2d.HydroRun(_params(nx, ny, impl, prec, **kw)) as run:
        run.init(ic)
        run.make_boundaries(0)
        chunks = [1, 6, steps // 2, steps - 7 - steps // 2]
        for c in chunks:
            run.run_steps(c)
        assert run.nStep == steps
        u, u2 = run.download(0), run.download(1)
        # dt of the next step = cfl / running max including the final state
        o = Oracle(nx, ny, bc=(kw["boundary_type_xmin"], kw["boundary_type_xmax"],
                               kw["boundary_type_ymin"], kw["boundary_type_ymax"]),
                   slope_type=kw["slope_type"], gamma0=kw["gamma0"],
                   precision="f32" if prec == e2d.PREC_FP32 else "f64")
        o.init(IC_DIAG if ic == "diag" else IC_IMPLODE)
        o.run(steps)
        assert run.get_dt() == o.compute_dt(steps % 2)
    assert _same(u, d["u_end"]), _diff_report(u, d["u_end"])
    assert _same(u2, d["u2_end"]), _diff_report(u2, d["u2_end"])


@pytest.mark.parametrize("impl", [1, 2])
def test_stage_arrays_step0(gpu_lib, golden, impl):
    """implementationVersions 1 and 2 keep q, fx, fy: equal to the reference's
    module arrays after step 0 (serial_ref/m_run.f90:14-17)."""
    d = np.load(golden / "run_diag128_fp64.npz")
    with e2d.HydroRun(_params(128, impl=impl)) as run:
        run.init("diag")
        dt = run.compute_dt(0)
        run.make_boundaries(0)
        dt = run.compute_dt(0)
        assert dt == d["dt"][0]
        run.godunov_unsplit(0, dt)
        for st in ("q", "fx", "fy"):
            a = run.download_stage(st)
            assert _same(a, d[f"stage0_{st}"]), f"{st}: {_diff_report(a, d['stage0_' + st])}"
        if impl == 2:
            with pytest.raises(e2d.E2DError, match="only the q, fx and fy"):
                run.download_stage("dqx")
        assert _same(run.download(1), d["snap_0"])


# ---------------------------------------------------------------------------
# other sizes against the oracle (strip and segment edges, minimal grids)

# 61 and 121 columns leave 1 output column to the last strip (packed in 16-lane
# groups), 80 leaves 20 (32-lane groups), 60, 7 and 4 make a single strip
SIZES = [(4, 4, (1, 1, 1, 1)), (61, 17, (1, 2, 3, 1)), (121, 33, (3, 3, 3, 3)),
         (60, 16, (2, 2, 2, 2)), (7, 130, (1, 3, 2, 1)), (80, 23, (3, 3, 2, 1)), (80, 41, (2, 1, 3, 3))]


@pytest.mark.parametrize("impl", [0, 1, 2])
@pytest.mark.parametrize("nx,ny,bc", SIZES)
def test_sizes_vs_oracle(gpu_lib, nx, ny, bc, impl):
    steps = 25
    kw = dict(boundary_type_xmin=bc[0], boundary_type_xmax=bc[1], boundary_type_ymin=bc[2],
              boundary_type_ymax=bc[3], slope_type=2)
    o = Oracle(nx, ny, bc=bc, slope_type=2)
    o.init(IC_IMPLODE)
    _, dts = o.run(steps)
    with e2d.HydroRun(_params(nx, ny, impl, **kw)) as run:
        run.init("implode")
        run.make_boundaries(0)
        run.run_steps(steps)
        u, u2 = run.download(0), run.download(1)
    assert _same(u, o.get(U)), _diff_report(u, o.get(U))
    assert _same(u2, o.get(U2)), _diff_report(u2, o.get(U2))


# segment plans of the fused step: long rows, short rows, percent of rows in
# long segments, strips per tile of the segment numbering (E2D_FUSED_SEG_ROWS /
# _SHORT_ROWS / _LONG_PCT / _TILE); None = default.  At 130 columns there are
# 3 strips: tile 2 leaves a last tile of one strip, 4 is clamped to 3.
# (E2D_FUSED_SEG_ROWS, E2D_FUSED_SHORT_ROWS, E2D_FUSED_LONG_PCT, E2D_FUSED_TILE, E2D_FUSED_PACK);
# at 130 columns the last strip has 10 output columns and is packed unless PACK is 0
PLANS = [(None, None, None, None, None), (5, 2, 60, None, None), (3, 1, 0, None, None), (7, 3, 100, None, None),
         (1, 1, 75, None, None), (40, 7, 90, None, None), (None, None, None, 2, None), (None, None, None, 3, None),
         (5, 2, 60, 4, None), (None, None, None, None, 0), (5, 2, 60, None, 0)]


@pytest.mark.parametrize("plan", PLANS)
@pytest.mark.parametrize("prec, kind", [(e2d.PREC_FP64, "f64"), (e2d.PREC_FP32, "f32")])
def test_segment_plans_vs_oracle(gpu_lib, monkeypatch, plan, prec, kind):
    """Whatever the split of each strip's rows into long and short segments,
    and however they are spread over the XCDs, the fused step gives the
    oracle's bits (device-resident dt loop, 20 steps, 130 x 151)."""
    for name, v in zip(("E2D_FUSED_SEG_ROWS", "E2D_FUSED_SHORT_ROWS", "E2D_FUSED_LONG_PCT", "E2D_FUSED_TILE",
                        "E2D_FUSED_PACK"), plan):
        if v is None:
            monkeypatch.delenv(name, raising=False)
        else:
            monkeypatch.setenv(name, str(v))
    nx, ny, steps = 130, 151, 20
    o = Oracle(nx, ny, precision=kind)
    o.init(IC_DIAG)
    _, dts = o.run(steps)
    o.compute_dt(steps % 2)  # the running maximum that includes the last state, as the device keeps it
    with e2d.HydroRun(_params(nx, ny, 0, prec)) as run:
        run.init("diag")
        run.make_boundaries(0)
        run.run_steps(steps)
        u, u2 = run.download(0), run.download(1)
        assert run.invdt == o.invdt
    assert _same(u, o.get(U)), _diff_report(u, o.get(U))
    assert _same(u2, o.get(U2)), _diff_report(u2, o.get(U2))


@pytest.mark.parametrize("prec", [e2d.PREC_FP32_NATIVE, e2d.PREC_MIXED])
@pytest.mark.parametrize("nx, ny", [(130, 151), (80, 64)])
def test_native_variants_independent_of_the_plan(gpu_lib, monkeypatch, prec, nx, ny):
    """fp32_native and mixed have no oracle of their own; their bits must not
    depend on how the fused step splits the grid: the default plan (last strip
    packed in 16- or 32-lane groups), the unpacked plan and a percentage plan
    with short segments give the same state after 20 steps."""
    out = []
    for env in ({}, {"E2D_FUSED_PACK": "0"}, {"E2D_FUSED_SEG_ROWS": "3", "E2D_FUSED_LONG_PCT": "50"}):
        for name in ("E2D_FUSED_PACK", "E2D_FUSED_SEG_ROWS", "E2D_FUSED_LONG_PCT"):
            monkeypatch.delenv(name, raising=False)
        for name, v in env.items():
            monkeypatch.setenv(name, v)
        with e2d.HydroRun(_params(nx, ny, 0, prec)) as run:
            run.init("diag")
            run.make_boundaries(0)
            run.run_steps(20)
            out.append((run.download(0), run.download(1), run.invdt))
    for u, u2, invdt in out[1:]:
        assert _same(u, out[0][0]) and _same(u2, out[0][1]) and invdt == out[0][2]


@pytest.mark.parametrize("impl", [0, 1, 2])
@pytest.mark.parametrize("prec, kind", [(e2d.PREC_FP64, "f64"), (e2d.PREC_FP32, "f32")])
def test_hllc_runs_vs_oracle(gpu_lib, impl, prec, kind):
    """riemannSolverType HLLC (riemann_2d_hllc, serial_ref/m_utils.f90:534-644)
    through whole runs: both initial conditions, mixed boundaries, minmod and
    MC slopes, all three implementation versions, against the oracle bit for
    bit (a NaN must sit where the oracle has one)."""
    for nx, ny, bc, st, ic, name in [(61, 17, (1, 2, 3, 1), 1, IC_DIAG, "diag"),
                                     (40, 52, (3, 3, 1, 2), 2, IC_IMPLODE, "implode")]:
        o = Oracle(nx, ny, bc=bc, slope_type=st, riemann_solver=2, precision=kind)
        o.init(ic)
        o.run(30)
        kw = dict(boundary_type_xmin=bc[0], boundary_type_xmax=bc[1], boundary_type_ymin=bc[2],
                  boundary_type_ymax=bc[3], slope_type=st, riemann_solver=e2d.RIEMANN_HLLC)
        with e2d.HydroRun(_params(nx, ny, impl, prec, **kw)) as run:
            run.init(name)
            run.make_boundaries(0)
            run.run_steps(30)
            u, u2 = run.download(0), run.download(1)
        assert _same(u, o.get(U)), f"{name}: " + _diff_report(u, o.get(U))
        assert _same(u2, o.get(U2)), f"{name}: " + _diff_report(u2, o.get(U2))


def test_boundaries_and_compute_dt_standalone(gpu_lib):
    nx, ny = 37, 29
    rng = np.random.default_rng(11)
    for bc in ((1, 1, 1, 1), (2, 2, 2, 2), (3, 3, 3, 3), (1, 2, 3, 2)):
        u = np.zeros((nx + 4, ny + 4, 4))
        u[..., 0] = rng.uniform(0.5, 2, u.shape[:2])
        u[..., 2] = rng.normal(0, 0.3, u.shape[:2])
        u[..., 3] = rng.normal(0, 0.3, u.shape[:2])
        u[..., 1] = rng.uniform(1, 3, u.shape[:2]) + 0.5 * (u[..., 2] ** 2 + u[..., 3] ** 2) / u[..., 0]
        o = Oracle(nx, ny, bc=bc)
        o.set(U, u)
        o.make_boundaries(0)
        kw = dict(boundary_type_xmin=bc[0], boundary_type_xmax=bc[1], boundary_type_ymin=bc[2],
                  boundary_type_ymax=bc[3])
        with e2d.HydroRun(_params(nx, ny, 1, **kw)) as run:
            run.upload(0, u)
            run.make_boundaries(0)
            assert _same(run.download(0), o.get(U))
            dt = run.compute_dt(0)
            assert dt == o.compute_dt(0)
            # running maximum: a calmer state does not raise dt
            run.upload(1, np.where(np.arange(4) == 0, u, u * 0.5))
            assert run.compute_dt(1) == dt


# ---------------------------------------------------------------------------
# the benchmark configuration: 2048^2 FP64

def test_2048_fp64_vs_oracle_first_steps(gpu_lib):
    """Bitwise against the oracle over the first steps at full size."""
    o = Oracle(2048)
    o.init(IC_DIAG)
    o.run(2)
    with e2d.HydroRun(_params(2048)) as run:
        run.init("diag")
        run.make_boundaries(0)
        run.run_steps(2)
        u = run.download(0)
        u2 = run.download(1)
    assert _same(u, o.get(U)) and _same(u2, o.get(U2))


def test_2048_fp64_fused_equals_materialised_and_symmetric(gpu_lib):
    """Size-independent properties after 200 steps at 2048^2: the three
    implementation versions agree bit for bit, and the diagonal problem stays
    exactly transpose-symmetric (rho = rho^T, rho*u = (rho*v)^T)."""
    res = []
    for impl in (0, 1, 2):
        with e2d.HydroRun(_params(2048, impl=impl)) as run:
            run.init("diag")
            run.make_boundaries(0)
            run.run_steps(200)
            res.append((run.download(0), run.get_dt(), run.sums(0)))
    a, dta, sa = res[0]
    for b, dtb, sb in res[1:]:
        assert dta == dtb
        assert _same(a, b), _diff_report(a, b)
        assert np.array_equal(sa, sb)
    i = slice(2, 2050)
    assert _same(a[i, i, 0], a[i, i, 0].T)
    assert _same(a[i, i, 2], a[i, i, 3].T)


def test_sums_match_host(gpu_lib):
    with e2d.HydroRun(_params(300, 200)) as run:
        run.init("implode")
        run.make_boundaries(0)
        run.run_steps(10)
        s = run.sums(0)
        u = run.download(0)
        again = run.sums(0)
    ref = u[2:302, 2:202, :].sum(axis=(0, 1))
    assert np.allclose(s, ref, rtol=1e-13, atol=0)
    assert np.array_equal(s, again)  # fixed summation order


# ---------------------------------------------------------------------------
# float variants without a reference kind: error against FP64 stated

# relative L1 after 200 steps of the 256^2 diagonal problem, per variable;
# measured on MI355X: fp32_native 1.07e-7 6.88e-8 3.02e-7 3.02e-7, mixed
# 4.15e-8 4.47e-8 1.71e-7 1.71e-7; bounds three times those
FLOAT_VARIANT_BOUNDS = {e2d.PREC_FP32_NATIVE: [3.2e-7, 2.1e-7, 9.1e-7, 9.1e-7],
                        e2d.PREC_MIXED: [1.25e-7, 1.35e-7, 5.2e-7, 5.2e-7]}


@pytest.mark.parametrize("prec", [e2d.PREC_FP32_NATIVE, e2d.PREC_MIXED])
def test_float_variants_against_fp64(gpu_lib, prec):
    out = {}
    for pr in (e2d.PREC_FP64, prec):
        with e2d.HydroRun(_params(256, prec=pr)) as run:
            run.init("diag")
            run.make_boundaries(0)
            run.run_steps(200)
            out[pr] = run.download(run.current())
    a, b = out[prec], out[e2d.PREC_FP64]
    i = slice(2, 258)
    errs = [rel_l1(a[i, i, v], b[i, i, v]) for v in range(4)]
    print(f"precision {prec} vs fp64, 256^2 diagonal problem, 200 steps: rel L1 {['%.3e' % e for e in errs]}")
    assert all(e <= bnd for e, bnd in zip(errs, FLOAT_VARIANT_BOUNDS[prec])), errs
