"""Benchmark of the euler2d hot path on one MI355X.

Default workload: 2048x2048 cells FP64, implementationVersion 0 (fused step),
the diagonal-discontinuity problem (x+y<1: rho=1,p=1 else rho=0.125,p=0.1),
gamma=1.4, CFL 0.8, minmod slopes, reflective walls.  A step is one pass of
the reference's main-loop body over the grid (compute_dt + godunov_unsplit,
serial_ref/main.f90:66-74); the state stays resident in HBM, dt stays on the
device.  Prints one JSON line.

  value     whole-job Mcell-updates/s = nx*ny*steps / wall time of the timed steps
  roofline  the fused step kernel against HBM: algorithmic bytes per launch
            (one read + one write of the 4 conserved variables per interior cell,
            64 B/cell in FP64) / average kernel duration from hipEvents recorded
            on the library's stream around each launch
  cpu_baseline  the reference serial solver (oracle/_ref, compiled from
            serial_ref/ with flang -O3, FP64 kind) on one core, 2048^2 for a few
            steps; the oracle's C restatement if that build is absent
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = ROOT / "profiles" / "traffic.json"


def _log(msg: str):
    print(msg, file=sys.stderr, flush=True)


def make_run(e2d, nx, impl, prec, problem="diag"):
    p = e2d.HydroParams(nx=nx, ny=nx, implementation_version=impl, precision=prec, problem=problem)
    run = e2d.HydroRun(p)
    run.init(problem, seed=20251202)
    run.make_boundaries(0)
    return run


def settle_clock(e2d, nx, impl, prec, ms: float) -> int:
    """Run steps of the same workload on a run of its own for `ms` ms of wall
    time, untimed, and free it; returns the step count.  An idle MI355X starts
    a launch series at a lower engine clock and raises it over the first tens
    of milliseconds.  Measured on the fused FP64 step at 2048^2: its
    GRBM_GUI_ACTIVE count stays at 3.6-3.7e6 cycles and its SQ_INSTS_VALU at
    8.76-8.82e7 per launch while the launch time falls from 208 to 185 us
    over the first 50 launches (2.23 to 2.46 GHz); back to back, launches
    take 202-214 us over the first 25 and 180-184 us from the 130th on.  The
    work per step is flat over those steps: the mean Newton count per face of
    the diagonal problem is 1.004-1.011 from step 5 to step 1000.  Settling
    first times the kernel at the clock a long run holds; the timed run
    itself still starts from the initial state."""
    if ms <= 0:
        return 0
    run = make_run(e2d, nx, impl, prec)
    n = 0
    try:
        t0 = time.perf_counter()
        while (time.perf_counter() - t0) * 1e3 < ms:
            run.run_steps(10)
            run.synchronize()
            n += 10
    finally:
        run.close()
    return n


def time_steps(run, steps: int, warmup: int) -> tuple[float, float]:
    """(seconds for `steps` timed steps, kernel-average ms from hipEvents)."""
    run.run_steps(warmup)
    run.synchronize()
    t0 = time.perf_counter()
    run.run_steps(steps)
    run.synchronize()
    wall = time.perf_counter() - t0
    # kernel durations, in a separate pass so event records do not perturb the timed one
    run.enable_timers(True)
    run.get_timers()
    run.run_steps(min(steps, 200))
    t = run.get_timers()
    run.enable_timers(False)
    ms, n = t["godunov"]
    return wall, ms / max(n, 1)


def cpu_baseline(nx: int, steps: int) -> dict:
    """Reference serial solver, one core, nx^2 for `steps` steps (bounded sample)."""
    ref = ROOT / "oracle" / "_ref" / "fp64" / "ref_driver"
    cores = 1
    if ref.exists():
        with tempfile.TemporaryDirectory() as td:
            nml = pathlib.Path(td) / "bench.nml"
            nml.write_text(
                f"&RUN\ntEnd=10.0\nnStepmax={steps}\nnOutput=100000\n/\n"
                f"&MESH\nnx={nx}\nny={nx}\nboundary_type_xmin=1\nboundary_type_xmax=1\n"
                "boundary_type_ymin=1\nboundary_type_ymax=1\n/\n"
                "&HYDRO\ngamma0=1.4\ncfl=0.8\nsmallr=1e-7\nsmallc=1e-7\nniter_riemann=10\n"
                "iorder=2\nslope_type=1\n/\n"
                "&ORACLE\ninitial_condition='diag'\ntiming=.true.\n/\n")
            cmd = [str(ref), str(nml)]
            if shutil_which("taskset"):
                cmd = ["taskset", "-c", "0"] + cmd
            out = subprocess.run(cmd, capture_output=True, text=True, check=True, timeout=600).stdout
        m = re.search(r"wall_s=\s*([0-9.Ee+-]+)\s+mcell_per_s=\s*([0-9.Ee+-]+)", out)
        return {"value": float(m.group(2)), "unit": "Mcell-updates/s", "cores": cores,
                "kind": "reference",
                "sample": f"{nx}x{nx} FP64, {steps} steps of the reference serial solver "
                          f"(serial_ref/ built with flang -O3, FP64 kind), wall {float(m.group(1)):.1f} s"}
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import IC_DIAG, Oracle, lib
    cores = lib().oracle_threads(1)
    o = Oracle(nx)
    o.init(IC_DIAG)
    t0 = time.perf_counter()
    o.run(steps)
    wall = time.perf_counter() - t0
    return {"value": nx * nx * steps / wall / 1e6, "unit": "Mcell-updates/s", "cores": cores,
            "kind": "port", "sample": f"{nx}x{nx} FP64, {steps} steps of the oracle C restatement, "
                                      f"wall {wall:.1f} s"}


def cpu_all_cores(nx: int, steps: int) -> dict:
    """The oracle C restatement on every host core (OpenMP over rows; the same
    bits as one core), nx^2 for `steps` steps: a second, separately reported
    CPU figure beside the single-core reference baseline."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import IC_DIAG, Oracle, lib
    threads = lib().oracle_threads(0)
    o = Oracle(nx)
    o.init(IC_DIAG)
    o.run(2)  # untimed: first touches of the arrays
    t0 = time.perf_counter()
    o.run(steps)
    wall = time.perf_counter() - t0
    lib().oracle_threads(1)
    return {"value": nx * nx * steps / wall / 1e6, "unit": "Mcell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{nx}x{nx} FP64, {steps} steps of the oracle C restatement on {threads} threads, "
                      f"wall {wall:.1f} s"}


def state_bytes(e2d, prec) -> int:
    """Bytes per stored conserved variable: 8 for the FP64 state of fp64 and
    mixed, 4 for the FP32 state of fp32 and fp32_native."""
    return 8 if prec in (e2d.PREC_FP64, e2d.PREC_MIXED) else 4


def shutil_which(name):
    import shutil
    return shutil.which(name)


def variant(e2d, nx, impl, prec, steps, warmup):
    run = make_run(e2d, nx, impl, prec)
    try:
        wall, kms = time_steps(run, steps, warmup)
    finally:
        run.close()
    # algorithmic HBM bytes per step: the 4 conserved variables read and
    # written once per cell, in the storage type (8 B for the FP64 state of
    # fp64 and mixed, 4 B for the FP32 state of fp32 and fp32_native)
    alg = nx * nx * 4 * 2 * state_bytes(e2d, prec)
    return {"mcell_per_s": round(nx * nx * steps / wall / 1e6, 1), "steps": steps,
            "ms_per_step": round(1e3 * wall / steps, 4), "kernel_ms": round(kms, 4),
            "hbm_frac_algorithmic": round(alg / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}


def host_loop(e2d, nx, steps, warmup=10):
    """The reference's loop shape: compute_dt (dt read back to the host) then
    godunov_unsplit, every step."""
    run = make_run(e2d, nx, 0, e2d.PREC_FP64)
    n = 0
    for _ in range(warmup):
        run.godunov_unsplit(n, run.compute_dt(n % 2))
        n += 1
    run.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run.godunov_unsplit(n, run.compute_dt(n % 2))
        n += 1
    run.synchronize()
    wall = time.perf_counter() - t0
    run.close()
    return {"mcell_per_s": round(nx * nx * steps / wall / 1e6, 1), "steps": steps,
            "ms_per_step": round(1e3 * wall / steps, 4)}


def precision_error(e2d, nx, steps):
    """FP32 variants against the FP64 run on the bumps problem: relative L1 and
    max-abs error per conserved variable after `steps` steps."""
    def final(prec):
        run = make_run(e2d, nx, 0, prec, problem="bumps")
        run.run_steps(steps)
        u = run.download(run.current())[2:nx + 2, 2:nx + 2, :]
        sums = run.sums(run.current())
        run.close()
        return u, sums

    ref, ref_sums = final(e2d.PREC_FP64)
    out = {}
    for name in ("fp32", "fp32_native", "mixed"):
        u, sums = final(e2d.PRECISIONS[name])
        d = u - ref
        del u
        out[name] = {
            "rel_l1": [float(np.abs(d[..., v]).sum() / np.abs(ref[..., v]).sum()) for v in range(4)],
            "linf": [float(np.abs(d[..., v]).max()) for v in range(4)],
            "mass_energy_sums_rel": [float(abs(sums[v] - ref_sums[v]) / abs(ref_sums[v])) for v in (0, 1)],
        }
    return out


def _rollout(e2d, nx, prec, dts, nsnap, lam_of, recorded=False):
    """Forward with the frozen dt sequence, then the checkpointed adjoint of
    lam_of(final state); returns (forward s, gradient s, recomputed steps,
    gradient array, final state).  recorded: the forward is
    e2d_rollout_forward, which keeps the reverse sweep's first checkpoints."""
    run = make_run(e2d, nx, 0, prec)
    u0 = run.download(0)
    run.synchronize()
    t0 = time.perf_counter()
    if recorded:
        run.rollout_forward(0, dts, nsnap)
    else:
        for n, d in enumerate(dts):
            run.godunov_unsplit(n, d)
    run.synchronize()
    t_fwd = time.perf_counter() - t0
    uT = run.download(len(dts) % 2)
    lam_n, lam_0 = run.field_alloc(), run.field_alloc()
    run.field_upload(lam_n, lam_of(uT))
    run.upload(0, u0)
    run.synchronize()
    t0 = time.perf_counter()
    nf = run.adjoint_rollout(0, dts, nsnap, lam_n, lam_0)
    t_adj = time.perf_counter() - t0
    grad = run.field_download(lam_0)
    run.close()
    return t_fwd, t_adj, nf, grad, uT


def adjoint_metrics(e2d):
    """Adjoint cost and accuracy.
    * one adjoint step (dense seeded lam) against one forward step at 2048^2,
      FP64, FP32 and mixed (kernel time from hipEvents, state after 100 steps);
    * 512^2 x 400 steps, 20 snapshots, J = sum of rho over [0.6,0.8]^2 (FP64);
    * 2048^2 x 2000 steps, 45 snapshots, J = sum over the interior of
      (rho_T - rho_target)^2 with a seeded smooth target: FP64 and FP32
      (forward and adjoint in float, lam accumulated in double), same frozen
      dt sequence (recorded from the FP64 run); the FP32 gradient's relative
      L2 error against the FP64 one; and the FP64 gradient again with 200
      snapshots (27 GB of HBM), which recomputes fewer forward steps, and
      after a recorded forward (e2d_rollout_forward), with 45 and 200.
    Forward times are the step launches alone (dt frozen), gradient times the
    whole checkpointed reverse sweep including recomputation."""
    out = {}
    for name, prec in e2d.PRECISIONS.items():
        run = make_run(e2d, 2048, 0, prec)
        run.run_steps(100)
        run.synchronize()
        buf, dt = run.current(), run.get_dt()
        a, b = run.field_alloc(), run.field_alloc()
        # a dense seeded lam: every face carries a nonzero flux bar (a zero bar
        # skips that face's Riemann adjoint)
        run.field_upload(a, np.random.default_rng(7).normal(size=(2052, 2052, 4)))
        run.adjoint_step(buf, dt, a, b)  # untimed: first launches load the kernels
        run.synchronize()
        run.enable_timers(True)
        run.get_timers()
        for _ in range(20):
            run.godunov_unsplit(run.nStep, dt)
        fwd = run.get_timers()["godunov"]
        for _ in range(20):
            run.adjoint_step(buf, dt, a, b)
        adj = run.get_timers()["godunov"]
        run.close()
        f_ms, a_ms = fwd[0] / fwd[1], adj[0] / adj[1]
        out[f"step_2048_{name}"] = {"forward_kernel_ms": round(f_ms, 4), "adjoint_kernels_ms": round(a_ms, 4),
                                    "adjoint_over_forward": round(a_ms / f_ms, 3)}

    def record_dts(nx, nsteps):
        run = make_run(e2d, nx, 0, e2d.PREC_FP64)
        dts = []
        for n in range(nsteps):
            d = run.compute_dt(n % 2)
            dts.append(d)
            run.godunov_unsplit(n, d)
        run.close()
        return dts

    nx = 512
    xc = (np.arange(nx + 4) - 2 + 0.5) / nx
    sel = (xc >= 0.6) & (xc <= 0.8)
    box = np.zeros((nx + 4, nx + 4, 4))
    box[np.ix_(sel, sel, [0])] = 1.0
    dts = record_dts(nx, 400)
    t_f, t_a, nf, _, _ = _rollout(e2d, nx, e2d.PREC_FP64, dts, 20, lambda uT: box)
    out["rollout_512_400steps_20snap_fp64"] = {"forward_s": round(t_f, 4), "gradient_s": round(t_a, 4),
                                               "gradient_over_forward": round(t_a / t_f, 3),
                                               "forward_steps_recomputed": nf}

    nx, nsteps, nsnap = 2048, 2000, 45
    X, Y = np.meshgrid((np.arange(nx + 4) - 1.5) / nx, (np.arange(nx + 4) - 1.5) / nx, indexing="ij")
    rng = np.random.default_rng(2048)
    k = rng.integers(1, 4, size=(3, 2))
    ph = rng.uniform(size=(3, 2))
    target = 1.0 + 0.3 * sum(np.sin(2 * np.pi * (k[m, 0] * X + ph[m, 0])) * np.cos(2 * np.pi * (k[m, 1] * Y + ph[m, 1]))
                             for m in range(3)) / 3
    inner = np.zeros((nx + 4, nx + 4), dtype=bool)
    inner[2:nx + 2, 2:nx + 2] = True

    def misfit_lam(uT):
        lam = np.zeros_like(uT)
        lam[..., 0] = np.where(inner, 2.0 * (uT[..., 0] - target), 0.0)
        return lam

    dts = record_dts(nx, nsteps)
    c5, grads = {}, {}
    for name, prec in (("fp64", e2d.PREC_FP64), ("fp32", e2d.PREC_FP32), ("mixed", e2d.PREC_MIXED)):
        t_f, t_a, nf, g, uT = _rollout(e2d, nx, prec, dts, nsnap, misfit_lam)
        grads[name] = g[2:nx + 2, 2:nx + 2, :]
        c5[name] = {"forward_s": round(t_f, 4), "gradient_s": round(t_a, 4),
                    "gradient_over_forward": round(t_a / t_f, 3), "forward_steps_recomputed": nf,
                    "J": float(np.sum(np.where(inner, (uT[..., 0] - target) ** 2, 0.0)))}
    # the same FP64 gradient with the snapshots sized for HBM rather than for
    # the configuration: 200 states (27 GB at 2048^2 FP64) cut the recomputation
    t_f, t_a, nf, g, _ = _rollout(e2d, nx, e2d.PREC_FP64, dts, 200, misfit_lam)
    c5["fp64_200_snapshots"] = {"gradient_s": round(t_a, 4), "gradient_over_forward": round(t_a / t_f, 3),
                                "forward_steps_recomputed": nf,
                                "same_gradient_as_45_snapshots": bool(np.array_equal(g[2:nx + 2, 2:nx + 2, :],
                                                                                     grads["fp64"]))}
    # the forward that keeps the reverse sweep's first checkpoints (what the
    # torch harness runs), 45 and 200 snapshots
    for ns in (45, 200):
        t_f, t_a, nf, g, _ = _rollout(e2d, nx, e2d.PREC_FP64, dts, ns, misfit_lam, recorded=True)
        c5[f"fp64_recorded_forward_{ns}_snapshots"] = {
            "forward_s": round(t_f, 4), "gradient_s": round(t_a, 4), "gradient_over_forward": round(t_a / t_f, 3),
            "forward_steps_recomputed": nf,
            "same_gradient_as_45_snapshots": bool(np.array_equal(g[2:nx + 2, 2:nx + 2, :], grads["fp64"]))}
    for name in ("fp32", "mixed"):
        d = grads[name] - grads["fp64"]
        c5[f"{name}_gradient_rel_l2_vs_fp64"] = float(np.linalg.norm(d) / np.linalg.norm(grads["fp64"]))
        c5[f"{name}_gradient_rel_l2_vs_fp64_per_variable"] = [
            float(np.linalg.norm(d[..., v]) / np.linalg.norm(grads["fp64"][..., v])) for v in range(4)]
    out["rollout_2048_2000steps_45snap_misfit"] = c5
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=2048)
    ap.add_argument("--impl", type=int, default=0)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "fp32_native", "mixed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--settle-ms", type=float, default=250.0,
                    help="before the warmup, run the workload untimed on a separate run for this many ms "
                         "so the engine clock has settled (0: off)")
    ap.add_argument("--adjoint-steps", type=int, default=0,
                    help="only time this many adjoint steps (dense seeded lam, state after 100 steps) "
                         "against as many forward steps at the workload size and exit")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (10 times) and exit; "
                         "its FETCH_SIZE/WRITE_SIZE counters calibrate the HBM byte counters")
    args = ap.parse_args(argv)

    if args.calibrate_copy_mib:
        import torch
        n = args.calibrate_copy_mib * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 10}))
        return 0

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the step runs on one GPU; it is not sharded across GPUs "
              "(a 2-D block decomposition with a 2-cell halo exchange would do that). Use --gpus 1.",
              file=sys.stderr)
        return 2

    import euler2d_cudafortran_amd as e2d
    e2d.load_library()
    if not e2d.device_available():
        print("bench.py: no HIP device", file=sys.stderr)
        return 1
    prec = e2d.PRECISIONS[args.precision]
    nx = args.nx

    if args.adjoint_steps:
        settle_clock(e2d, nx, 0, prec, args.settle_ms)
        run = make_run(e2d, nx, 0, prec)
        run.run_steps(100)
        run.synchronize()
        buf, dt = run.current(), run.get_dt()
        a, b = run.field_alloc(), run.field_alloc()
        run.field_upload(a, np.random.default_rng(7).normal(size=(nx + 4, nx + 4, 4)))
        run.adjoint_step(buf, dt, a, b)  # untimed: first launches load the kernels
        run.synchronize()
        run.enable_timers(True)
        run.get_timers()
        for _ in range(args.adjoint_steps):
            run.adjoint_step(buf, dt, a, b)
        adj = run.get_timers()["godunov"]
        for _ in range(args.adjoint_steps):
            run.godunov_unsplit(run.nStep, dt)
        fwd = run.get_timers()["godunov"]
        run.close()
        a_ms, f_ms = adj[0] / adj[1], fwd[0] / fwd[1]
        print(json.dumps({"metric": f"adjoint step / forward step at {nx}^2 {args.precision}", "value": round(a_ms / f_ms, 3),
                          "adjoint_step_ms": round(a_ms, 4), "forward_step_ms": round(f_ms, 4),
                          "steps": args.adjoint_steps, "higher_is_better": False}))
        return 0

    _log(f"[bench] {nx}x{nx} {args.precision} impl {args.impl}: {args.warmup} warmup + {args.steps} steps")
    run = make_run(e2d, nx, args.impl, prec)
    settle_steps = settle_clock(e2d, nx, args.impl, prec, args.settle_ms)
    wall, kernel_ms = time_steps(run, args.steps, args.warmup)
    dt_final = run.get_dt()
    # copy ceiling: device-to-device copies of one state buffer through the
    # C-ABI, host-timed (each call synchronises, so this is a lower estimate)
    dst = run.field_alloc()
    run.copy_from_buffer(0, dst)
    reps = 50
    t0 = time.perf_counter()
    for _ in range(reps):
        run.copy_from_buffer(0, dst)
    copy_s = (time.perf_counter() - t0) / reps
    buf_bytes = run.device_buffer(0)[1] * (nx + 4) * 4 * state_bytes(e2d, prec)
    copy_gbs = 2 * buf_bytes / copy_s / 1e9
    run.close()
    cells = nx * nx
    value = cells * args.steps / wall / 1e6
    ms_per_step = 1e3 * wall / args.steps
    alg_bytes = cells * 4 * state_bytes(e2d, prec) * 2  # read + write of 4 conserved variables
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    traffic, valu = None, None
    if TRAFFIC_FILE.exists():
        tr = json.loads(TRAFFIC_FILE.read_text())
        key = f"{nx}_{args.precision}_impl{args.impl}"
        if key in tr:
            traffic = tr[key]["bytes_per_launch"]
            if "valu_issue_us" in tr[key]:
                # VALU issue time per launch from the counter run, against this run's kernel time
                valu = {"bound": tr[key].get("valu_bound", "valu issue"), "issue_us": tr[key]["valu_issue_us"],
                        "kernel_us": round(kernel_ms * 1e3, 1),
                        "frac": round(tr[key]["valu_issue_us"] / (kernel_ms * 1e3), 3),
                        "lane_instructions_per_cell": tr[key]["valu_lane_instructions_per_cell"],
                        "source": tr[key]["valu_source"]}
    _log(f"[bench] {value:.1f} Mcell/s, {ms_per_step:.4f} ms/step, kernel {kernel_ms:.4f} ms, "
         f"{achieved:.1f} GB/s algorithmic")

    variants = {}
    if not args.no_variants:
        for name, (vnx, impl, vprec, vsteps) in {
            "fp64_materialised_2048": (2048, 1, e2d.PREC_FP64, 200),
            "fp64_flux_store_2048": (2048, 2, e2d.PREC_FP64, 200),
            "fp32_fused_2048": (2048, 0, e2d.PREC_FP32, 500),
            "fp32_native_fused_2048": (2048, 0, e2d.PREC_FP32_NATIVE, 500),
            "mixed_fused_2048": (2048, 0, e2d.PREC_MIXED, 500),
            "fp64_fused_4096": (4096, 0, e2d.PREC_FP64, 200),
            "fp32_fused_4096": (4096, 0, e2d.PREC_FP32, 500),
            "fp32_native_fused_4096": (4096, 0, e2d.PREC_FP32_NATIVE, 500),
            "mixed_fused_4096": (4096, 0, e2d.PREC_MIXED, 500),
        }.items():
            _log(f"[bench] variant {name}")
            variants[name] = variant(e2d, vnx, impl, vprec, vsteps, 10)
        _log("[bench] variant fp64_host_loop_2048")
        variants["fp64_host_loop_2048"] = host_loop(e2d, 2048, 200)
        _log("[bench] precision error, bumps problem 4096^2, 500 steps")
        variants["precision_error_bumps_4096_500steps_vs_fp64"] = precision_error(e2d, 4096, 500)

    adjoint = None
    if not args.no_variants:
        _log("[bench] adjoint cost")
        adjoint = adjoint_metrics(e2d)

    cpu = cpu_all = None
    if not args.no_cpu_baseline:
        _log(f"[bench] cpu baseline {nx}x{nx}, {args.cpu_steps} steps")
        cpu = cpu_baseline(nx, args.cpu_steps)
        _log(f"[bench] oracle on all host cores {nx}x{nx}, 20 steps")
        cpu_all = cpu_all_cores(nx, 20)

    line = {
        "metric": "Mcell-updates/s at 2048^2 FP64 vs FP32 (%HBM peak); adjoint/forward cost ratio",
        "value": round(value, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        # untimed steps of a separate run before the warmup (settle_clock)
        "clock_settle": {"ms": args.settle_ms, "steps": settle_steps},
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "fp32_native": "f32", "mixed": "f32+f64"}[args.precision],
        "data": "synthetic",
        "config": {"workload": f"euler2d {nx}x{nx} {args.precision}, implementationVersion {args.impl}, "
                               "diagonal discontinuity, gamma 1.4, CFL 0.8, minmod, reflective walls",
                   "nx": nx, "ny": nx, "precision": args.precision,
                   "implementation_version": args.impl,
                   "mcell_updates_per_s_reference_formula": round((nx + 4) ** 2 * args.steps / wall / 1e6, 2),
                   "dt_after_run": dt_final},
        # "bound" names the roofline that achieved / peak / frac are stated
        # against (HBM bytes, the stencil's natural roof); "binding" names the
        # limit that holds in fact, with its issue fraction (VALU issue time
        # from the counter run over this run's kernel time) when one is known
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "binding": valu["bound"] if valu else "hbm",
                     "valu_issue_frac": valu["frac"] if valu else None,
                     "limiter": valu,
                     "kernel": {0: "k_step_fused", 1: "materialised step (5 kernels)",
                                2: "flux-store step (3 kernels)"}[args.impl],
                     "kernel_avg_ms": round(kernel_ms, 5),
                     "copy_ceiling_gbs": round(copy_gbs, 1),
                     "algorithmic_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
        "cpu_all_cores": cpu_all,
        "variants": variants,
        "adjoint": adjoint,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
