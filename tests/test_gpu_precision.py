"""The precision variants on the seeded bumps problem.

* The device's initial state of the bumps problem equals the host
  restatement (tests/bumps_ref.py: SplitMix64 draws, Gaussian sums) at the
  configuration's 4096^2 in FP64, and at 256^2 in every precision.
* Configuration C3: 4096^2, diagonal discontinuity plus 16 seeded Gaussian
  bumps, 500 steps of the device-resident dt loop, implementationVersion 0.
  The FP64 run is the reference (pinned to the oracle bit for bit by
  test_gpu_long_parity.py over the first 100 steps of the same problem); the
  float variants are held to per-variable relative-L1 and max-abs error bounds
  set from their measured errors with a stated margin:
    fp32         FP32 state and arithmetic (the reference's default kind)
    fp32_native  FP32 state, native reciprocal / square root, FP64 update add
    mixed        FP64 state and update, FP32 slope / trace / Riemann arithmetic
                 (native reciprocal / square root)
  and the interior mass and energy sums (FP64, fixed order) against FP64's.
"""
import numpy as np
import pytest

import euler2d_cudafortran_amd as e2d
from bumps_ref import bumps_state

pytestmark = pytest.mark.gpu

SEED = 20251202
NAMES = {e2d.PREC_FP64: "fp64", e2d.PREC_FP32: "fp32", e2d.PREC_FP32_NATIVE: "fp32_native",
         e2d.PREC_MIXED: "mixed"}
STORE = {e2d.PREC_FP64: np.float64, e2d.PREC_FP32: np.float32, e2d.PREC_FP32_NATIVE: np.float32,
         e2d.PREC_MIXED: np.float64}
ARITH = {e2d.PREC_FP64: np.float64, e2d.PREC_FP32: np.float32, e2d.PREC_FP32_NATIVE: np.float64,
         e2d.PREC_MIXED: np.float64}

# C3 bounds after 500 steps at 4096^2: per variable (rho, E, rho*u, rho*v)
# relative L1 and max-abs error against the FP64 run, and the relative error
# of the interior mass and energy sums.  Measured (MI355X):
#   fp32         rel-L1 4.15e-7 1.31e-7 1.00e-6 1.04e-6   max-abs 2.02e-5 1.55e-4 1.16e-5 5.50e-5
#                mass 5.4e-8  energy 4.4e-8
#   fp32_native  rel-L1 4.10e-7 1.17e-7 9.56e-7 1.01e-6   max-abs 1.72e-5 1.31e-4 1.00e-5 4.67e-5
#                mass 5.3e-8  energy 1.9e-8
#   mixed        rel-L1 4.18e-8 5.03e-8 4.28e-7 4.31e-7   max-abs 1.07e-6 7.54e-6 2.34e-6 2.65e-6
#                mass 2.0e-16 energy 9.9e-15
# bounds: three times the measured values
BOUNDS = {
    "fp32": {"rel_l1": [1.25e-6, 4e-7, 3e-6, 3.1e-6], "linf": [6.1e-5, 4.7e-4, 3.5e-5, 1.65e-4], "sums": 1.6e-7},
    "fp32_native": {"rel_l1": [1.23e-6, 3.5e-7, 2.9e-6, 3e-6], "linf": [5.2e-5, 3.9e-4, 3e-5, 1.4e-4],
                    "sums": 1.6e-7},
    "mixed": {"rel_l1": [1.3e-7, 1.5e-7, 1.3e-6, 1.3e-6], "linf": [3.2e-6, 2.3e-5, 7e-6, 8e-6], "sums": 3e-14},
}


def _run(nx, prec, steps):
    p = e2d.HydroParams(nx=nx, ny=nx, precision=prec, problem="bumps")
    with e2d.HydroRun(p) as run:
        run.init("bumps", seed=SEED)
        u0 = run.download(0)
        run.make_boundaries(0)
        if steps:
            run.run_steps(steps)
        u = run.download(run.current())
        sums = run.sums(run.current())
    return u0, u, sums


def _check_initial(u0, nx, prec):
    ref = bumps_state(nx, nx, 1.4, SEED, arith=ARITH[prec], store=STORE[prec])
    # the device's exp and the host's may differ in the last bit of a bump
    # weight: a few units in the last place of the storage type
    ulp = np.finfo(STORE[prec]).eps
    for v in (0, 1):
        rel = np.abs(u0[..., v] - ref[..., v]) / np.abs(ref[..., v])
        print(f"{NAMES[prec]} {nx}^2 initial state, variable {v}: max relative difference {rel.max():.2e}, "
              f"{np.count_nonzero(rel)} cells differ", flush=True)
        assert rel.max() <= 4 * ulp
    assert np.all(u0[..., 2:] == 0)


@pytest.mark.parametrize("prec", [e2d.PREC_FP64, e2d.PREC_FP32, e2d.PREC_FP32_NATIVE, e2d.PREC_MIXED])
def test_bumps_initial_state_256(gpu_lib, prec):
    u0, _, _ = _run(256, prec, 0)
    _check_initial(u0, 256, prec)


def test_c3_4096_bumps_500_steps(gpu_lib):
    nx, steps = 4096, 500
    i = slice(2, nx + 2)
    print(f"C3: {nx}^2 bumps, {steps} steps, precision fp64", flush=True)
    u0, ref, ref_sums = _run(nx, e2d.PREC_FP64, steps)
    _check_initial(u0, nx, e2d.PREC_FP64)
    del u0
    ref = ref[i, i, :]
    failures = []
    for prec in (e2d.PREC_FP32, e2d.PREC_FP32_NATIVE, e2d.PREC_MIXED):
        name = NAMES[prec]
        print(f"C3: {nx}^2 bumps, {steps} steps, precision {name}", flush=True)
        _, u, sums = _run(nx, prec, steps)
        d = u[i, i, :] - ref
        del u
        rel_l1 = [float(np.abs(d[..., v]).sum() / np.abs(ref[..., v]).sum()) for v in range(4)]
        linf = [float(np.abs(d[..., v]).max()) for v in range(4)]
        mass_rel = abs(sums[0] - ref_sums[0]) / abs(ref_sums[0])
        energy_rel = abs(sums[1] - ref_sums[1]) / abs(ref_sums[1])
        print(f"C3 {name}: rel-L1 {['%.2e' % x for x in rel_l1]} max-abs {['%.2e' % x for x in linf]} "
              f"mass {mass_rel:.2e} energy {energy_rel:.2e}", flush=True)
        b = BOUNDS[name]
        for v in range(4):
            if rel_l1[v] > b["rel_l1"][v] or linf[v] > b["linf"][v]:
                failures.append((name, v, rel_l1[v], linf[v]))
        if mass_rel > b["sums"] or energy_rel > b["sums"]:
            failures.append((name, "sums", mass_rel, energy_rel))
    assert not failures, failures
