"""Check of the torch.autograd harness (euler2d_cudafortran_amd/autograd.py),
run by tests/test_gpu_torch.py in its own process: torch must load its HIP
runtime before libeuler2d_hip does (both carry the soname libamdhip64.so.7),
so the harness is exercised in a fresh interpreter that imports torch first.

Objective: J = sum over the interior of (rho_N - rho_target)^2 after the
rollout (rho_target a smooth seeded field), so autograd hands the adjoint a
non-trivial output gradient 2 (rho_N - rho_target).  The harness gradient must
equal the C-ABI rollout adjoint fed with that same output gradient, bit for
bit, and match the directional derivative of J along a smooth (rho, p)
direction from a smooth initial state with flow, dJ = sum 2 (rho_N -
rho_target) d rho_N, with d rho_N the 4-point finite difference of the rollout
evaluated by the oracle in binary128 at eps = 1e-10 with the frozen dt
sequence and every decision of the step frozen at the unperturbed run's (no
truncation or roundoff that a double shows; bar 1e-10).  Prints one line "AUTOGRAD_CHECK ok rel=<relative error>"
on success."""
import pathlib
import sys

import torch  # noqa: I001  (before the library: see above)
import numpy as np

sys.path.insert(0, str(pathlib.Path(__file__).resolve().parents[2]))
sys.path.insert(0, str(pathlib.Path(__file__).resolve().parents[1]))
import euler2d_cudafortran_amd as e2d  # noqa: E402
from oracle_lib import Oracle  # noqa: E402


def _smooth_field(rng, X, Y, modes=4):
    f = np.zeros_like(X)
    for _ in range(modes):
        kx, ky = rng.integers(1, 4, 2)
        f += rng.normal() * np.sin(2 * np.pi * (kx * X + rng.uniform())) * np.cos(2 * np.pi * (ky * Y + rng.uniform()))
    return f / modes


def main(niter=10, eps=1e-10, tol=1e-10):
    from euler2d_cudafortran_amd.autograd import GodunovRollout, interior_mask, record_dts, state_tensor

    assert torch.cuda.is_available()

    nx, nsteps, gamma = 64, 40, 1.4
    with e2d.HydroRun(e2d.HydroParams(nx=nx, ny=nx, boundary_type_xmin=3, boundary_type_xmax=3,
                                      boundary_type_ymin=3, boundary_type_ymax=3,
                                      niter_riemann=niter)) as run:
        run.init("diag")
        base = state_tensor(run, 0)
        # device layout (var, j, i): rows are y, columns x
        J_, I_ = np.meshgrid(np.arange(base.shape[1]), np.arange(base.shape[2]), indexing="ij")
        x, y = (I_ - 1.5) / nx, (J_ - 1.5) / nx
        rng = np.random.default_rng(31)
        rho = 1 + 0.3 * _smooth_field(rng, x, y)
        u = 0.2 * _smooth_field(rng, x, y)
        v = 0.2 * _smooth_field(rng, x, y)
        p = 1 + 0.3 * _smooth_field(rng, x, y)
        u0 = base.clone()
        u0[0] = torch.from_numpy(rho)
        u0[1] = torch.from_numpy(p / (gamma - 1) + 0.5 * rho * (u * u + v * v))
        u0[2] = torch.from_numpy(rho * u)
        u0[3] = torch.from_numpy(rho * v)
        target = torch.from_numpy(1 + 0.3 * _smooth_field(rng, x, y)).cuda()
        dts = record_dts(run, u0, nsteps)
        inner = interior_mask(run, u0)[0]

        def misfit(uN):
            return torch.where(inner, (uN[0] - target) ** 2, torch.zeros_like(target)).sum()

        u0g = u0.clone().requires_grad_(True)
        uN = GodunovRollout.apply(u0g, run, dts, 6)
        J = misfit(uN)
        J.backward()
        grad = u0g.grad

        # the same gradient straight from the C-ABI
        lam_n = torch.zeros_like(u0)
        lam_n[0] = torch.where(inner, 2 * (uN.detach()[0] - target), torch.zeros_like(target))
        lam_0 = torch.zeros_like(lam_n)
        torch.cuda.synchronize()
        run.copy_to_buffer(0, u0.data_ptr())
        run.adjoint_rollout(0, dts, 3, lam_n.data_ptr(), lam_0.data_ptr())
        lam_0 = torch.where(interior_mask(run, lam_0), lam_0, torch.zeros_like(lam_0))
        assert torch.equal(grad, lam_0)

        # directional derivative in (rho, p) at fixed velocity against 4-point finite differences
        drho = torch.from_numpy(_smooth_field(rng, x, y)).cuda()
        dp = torch.from_numpy(_smooth_field(rng, x, y)).cuda()
        uu, vv = torch.from_numpy(u).cuda(), torch.from_numpy(v).cuda()
        du = torch.zeros_like(u0)
        du[0] = drho
        du[1] = dp / (gamma - 1) + 0.5 * drho * (uu * uu + vv * vv)
        du[2] = drho * uu
        du[3] = drho * vv
        du = torch.where(interior_mask(run, du), du, torch.zeros_like(du))
        ad = float((grad * du).sum())
        weight = torch.zeros_like(u0)
        weight[0] = torch.where(inner, 2 * (uN.detach()[0] - target), torch.zeros_like(target))

    def to_host(t):
        # device layout (4, jsize, pitch) -> the oracle's (isize, jsize, 4)
        return np.ascontiguousarray(t.detach().cpu().numpy()[:, :, :nx + 4].transpose(2, 1, 0))

    bc = (3, 3, 3, 3)
    o = Oracle(nx, nx, bc=bc, niter_riemann=niter, precision="f128")
    fd = o.fd(to_host(u0), to_host(du), dts, eps, weight=to_host(weight), freeze=True)
    rel = abs(fd - ad) / abs(fd)
    print(f"autograd gradient: adjoint {ad:.15e} binary128 finite difference {fd:.15e} rel {rel:.2e}")
    assert rel <= tol, rel
    print(f"AUTOGRAD_CHECK ok rel={rel:.3e}", flush=True)


if __name__ == "__main__":
    main()
