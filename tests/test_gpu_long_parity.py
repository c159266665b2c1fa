"""The benchmark configuration end to end: 2048^2, diagonal-discontinuity
problem, reflective walls, 1000 steps, FP64 and FP32 (the reference's default
kind), implementationVersion 0, 1 and 2 on the GPU (device-resident dt loop),
against the oracle of the same kind on every host core (its grid
loops run rows in parallel, which changes no bit: each cell is computed alone
and the dt reduction is a max).  The reference needs about 35 minutes of one
core for this run.

Bar: the stated tolerance rel-L1 <= 1e-12 per conserved variable, and in fact
bit identity (a zero may differ in sign) of both state buffers, ghost cells and
corners included, and of the running invDt and the dt of the next step.

On a 16-core host the FP64 oracle takes about a minute (52 s measured) and the
GPU 1.4 s for the three versions; E2D_LONG_PARITY=0 skips it.
"""
import os
import time

import numpy as np
import pytest

import euler2d_cudafortran_amd as e2d
from oracle_lib import IC_DIAG, U, U2, Oracle, lib, rel_l1

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(os.environ.get("E2D_LONG_PARITY") == "0",
                                 reason="E2D_LONG_PARITY=0 skips the long run")]

NX, STEPS = 2048, 1000


def _same(a, b):
    na, nb = np.isnan(a), np.isnan(b)
    return bool(np.array_equal(na, nb) and np.all(a[~na] == b[~nb]))


@pytest.mark.parametrize("prec, kind", [(e2d.PREC_FP64, "f64"), (e2d.PREC_FP32, "f32")])
def test_2048_1000_steps_vs_oracle(gpu_lib, prec, kind):
    gpu = {}
    for impl in (0, 1, 2):
        p = e2d.HydroParams(nx=NX, ny=NX, implementation_version=impl, precision=prec)
        with e2d.HydroRun(p) as run:
            run.init("diag")
            run.make_boundaries(0)
            t0 = time.perf_counter()
            run.run_steps(STEPS)
            run.synchronize()
            wall = time.perf_counter() - t0
            gpu[impl] = (run.download(0), run.download(1), run.get_dt(), run.invdt)
        print(f"GPU implementationVersion {impl}: {STEPS} steps in {wall:.2f} s", flush=True)

    threads = lib().oracle_threads(0)
    o = Oracle(NX, precision=kind)
    o.init(IC_DIAG)
    t0 = time.perf_counter()
    # the reference main loop (serial_ref/main.f90:36-76), as oracle_run runs it
    o.compute_dt(0)
    o.make_boundaries(0)
    for n in range(STEPS):
        dt = o.compute_dt(n % 2)
        o.godunov(n, dt)
        if (n + 1) % 100 == 0:
            print(f"oracle step {n + 1}/{STEPS}: {time.perf_counter() - t0:.0f} s on {threads} threads", flush=True)
    ref_u, ref_u2 = o.get(U), o.get(U2)
    next_dt = o.compute_dt(STEPS % 2)  # the dt the next step would use, as e2d_get_dt reports it
    ref_invdt = o.invdt

    for impl, (u, u2, gdt, ginvdt) in gpu.items():
        errs = [rel_l1(u[2:-2, 2:-2, v], ref_u[2:-2, 2:-2, v]) for v in range(4)]
        print(f"implementationVersion {impl}: rel-L1 per variable {errs}", flush=True)
        assert max(errs) <= 1e-12
        assert gdt == next_dt and ginvdt == ref_invdt, (impl, gdt, next_dt, ginvdt, ref_invdt)
        assert _same(u, ref_u) and _same(u2, ref_u2), f"implementationVersion {impl} differs in some bit"


def test_4096_bumps_fp64_vs_oracle(gpu_lib):
    """The precision variants' 4096^2 problem (diagonal discontinuity plus 16
    seeded Gaussian bumps) in FP64, 100 steps of the device-resident dt loop:
    at 4096^2 the fused step's long segments fill two rounds of resident
    wavefronts (one at 2048^2).  The oracle starts from the GPU's initial
    state (it has no bumps initialiser); bit identity as above."""
    nx, steps = 4096, 100
    p = e2d.HydroParams(nx=nx, ny=nx, precision=e2d.PREC_FP64)
    with e2d.HydroRun(p) as run:
        run.init("bumps")
        u0, u20 = run.download(0), run.download(1)
        run.make_boundaries(0)
        run.run_steps(steps)
        run.synchronize()
        u, u2, gdt, ginvdt = run.download(0), run.download(1), run.get_dt(), run.invdt
    o = Oracle(nx, precision="f64")
    o.init(IC_DIAG)
    o.set(U, u0)
    o.set(U2, u20)
    t0 = time.perf_counter()
    o.compute_dt(0)
    o.make_boundaries(0)
    for n in range(steps):
        dt = o.compute_dt(n % 2)
        o.godunov(n, dt)
    print(f"oracle {steps} steps at {nx}^2: {time.perf_counter() - t0:.0f} s", flush=True)
    ref_u, ref_u2 = o.get(U), o.get(U2)
    next_dt = o.compute_dt(steps % 2)
    assert gdt == next_dt and ginvdt == o.invdt, (gdt, next_dt, ginvdt, o.invdt)
    assert _same(u, ref_u) and _same(u2, ref_u2), "4096^2 bumps run differs in some bit"


@pytest.mark.parametrize("prec, kind", [(e2d.PREC_FP64, "f64"), (e2d.PREC_FP32, "f32")])
def test_wide_grid_packed_last_strip_vs_oracle(gpu_lib, prec, kind):
    """A large non-square grid whose last strip is packed in 32-lane groups
    (4100 columns: 68 full strips and 20 columns left) and whose FP64 long
    segments fill two rounds of resident wavefronts: 4100 x 2500, diagonal
    problem, 20 steps of the device-resident dt loop, bit identity as above."""
    nx, ny, steps = 4100, 2500, 20
    p = e2d.HydroParams(nx=nx, ny=ny, precision=prec)
    with e2d.HydroRun(p) as run:
        run.init("diag")
        run.make_boundaries(0)
        run.run_steps(steps)
        run.synchronize()
        u, u2, gdt, ginvdt = run.download(0), run.download(1), run.get_dt(), run.invdt
    o = Oracle(nx, ny, precision=kind)
    o.init(IC_DIAG)
    o.compute_dt(0)
    o.make_boundaries(0)
    for n in range(steps):
        dt = o.compute_dt(n % 2)
        o.godunov(n, dt)
    ref_u, ref_u2 = o.get(U), o.get(U2)
    next_dt = o.compute_dt(steps % 2)
    assert gdt == next_dt and ginvdt == o.invdt, (gdt, next_dt, ginvdt, o.invdt)
    assert _same(u, ref_u) and _same(u2, ref_u2), f"{nx} x {ny} {kind} run differs in some bit"
