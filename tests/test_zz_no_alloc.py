"""The allocation-free step loop (e2d_run_steps_noalloc) against e2d_run_steps.

e2d_run_steps_noalloc runs the device-resident step loop on caller-owned
buffers with its two running-invDt slots in a static device array of the
library.  Here the caller's buffers are the state buffers of a second context
(e2d_device_buffer) and the running invDt lives in the first double of a field
(e2d_field_alloc), so the test itself allocates through the C-ABI only.  Bar:
both state buffers (ghost cells and corners included) and the running invDt
bit-identical to e2d_run_steps from the same state, on two calls in a row (the
second on a smaller grid, so the shared scratch must be cleared by the call),
plus a continued call that starts from a nonzero running invDt.
"""
import ctypes

import numpy as np
import pytest

import euler2d_cudafortran_amd as e2d

pytestmark = pytest.mark.gpu


def _hip_sync():
    # the null stream does not order the contexts' non-blocking streams: wait for the device
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipDeviceSynchronize.restype = ctypes.c_int
    assert hip.hipDeviceSynchronize() == 0


def _same(a, b):
    na, nb = np.isnan(a), np.isnan(b)
    return bool(np.array_equal(na, nb) and np.all(a[~na] == b[~nb]))


def _params(nx, ny, prec, **kw):
    p = e2d.HydroParams(nx=nx, ny=ny, precision=prec, problem="diag")
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _reference(p, chunks):
    """e2d_run_steps over the chunks of steps: the final buffers and running invDt."""
    with e2d.HydroRun(p) as run:
        run.init("diag")
        run.make_boundaries(0)
        for k in chunks:
            run.run_steps(k)
        return run.download(0), run.download(1), run.invdt


def _noalloc(p, chunks):
    """The same steps through e2d_run_steps_noalloc on a context's buffers."""
    with e2d.HydroRun(p) as run:
        run.init("diag")
        run.make_boundaries(0)
        d_u, pitch = run.device_buffer(0)
        d_u2, _ = run.device_buffer(1)
        d_invdt = run.field_alloc()  # zero-filled: the running invDt of a fresh run
        run.synchronize()
        n = 0
        for k in chunks:
            print(f"e2d_run_steps_noalloc {p.nx}x{p.ny} precision {p.precision}: {k} steps from step {n}",
                  flush=True)
            e2d.run_steps_noalloc(p, d_u, d_u2, pitch, k, n, d_invdt, 0)
            n += k
        _hip_sync()
        invdt = run.field_download(d_invdt)[0, 0, 0]
        return run.download(0), run.download(1), invdt


@pytest.mark.parametrize("prec", [e2d.PREC_FP64, e2d.PREC_FP32, e2d.PREC_MIXED])
def test_noalloc_matches_run_steps(gpu_lib, prec):
    # first call on a larger grid, second on a smaller one (the scratch keeps
    # the first call's maxima, which exceed nothing only if cleared)
    for nx, ny, chunks in ((96, 80, (23,)), (40, 36, (17,))):
        p = _params(nx, ny, prec)
        ref = _reference(p, chunks)
        got = _noalloc(p, chunks)
        assert _same(got[0], ref[0]) and _same(got[1], ref[1]), (nx, ny, prec)
        assert got[2] == ref[2], (got[2], ref[2])


def test_noalloc_continues_running_invdt(gpu_lib):
    """Steps in three calls (odd counts, so the input buffer alternates) equal one run_steps loop."""
    p = _params(64, 48, e2d.PREC_FP64, boundary_type_xmax=2, boundary_type_ymin=3, boundary_type_ymax=3)
    ref = _reference(p, (5, 8, 7))
    got = _noalloc(p, (5, 8, 7))
    assert _same(got[0], ref[0]) and _same(got[1], ref[1])
    assert got[2] == ref[2]


def test_noalloc_rejects_bad_arguments(gpu_lib):
    lib = e2d.load_library()
    p = _params(32, 32, e2d.PREC_FP64)
    with e2d.HydroRun(p) as run:
        d_u, pitch = run.device_buffer(0)
        d_u2, _ = run.device_buffer(1)
        d_invdt = run.field_alloc()
        for kw, args, msg in (
                (dict(implementation_version=1), (d_u, d_u2, pitch), "implementationVersion 0"),
                ({}, (d_u, d_u, pitch), "distinct"),
                ({}, (d_u, d_u2, 35), "pitch"),
                ({}, (None, d_u2, pitch), "NULL")):
            q = _params(32, 32, e2d.PREC_FP64, **kw)
            with pytest.raises(e2d.E2DError, match=msg):
                e2d.run_steps_noalloc(q, *args, 1, 0, d_invdt)
        with pytest.raises(e2d.E2DError, match="nsteps"):
            e2d.run_steps_noalloc(p, d_u, d_u2, pitch, -1, 0, d_invdt)
        assert lib.e2d_run_steps_noalloc(ctypes.byref(p.to_c()), d_u, d_u2, pitch, 0, 0, d_invdt, None) == 0
