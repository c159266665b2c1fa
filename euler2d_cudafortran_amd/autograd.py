"""torch.autograd harness over the adjoint C-ABI (gradient checks, inverse problems).

`GodunovRollout.apply(u0, run, dts, nsnap)` maps the initial state u0 (FP64
CUDA tensor of shape (4, jsize, pitch): the library's device layout, see
`state_tensor`) to the state after len(dts) steps with the frozen dt
sequence dts.  Its forward is e2d_rollout_forward, which keeps the first
checkpoints of the reverse sweep, and its backward e2d_adjoint_rollout
(binomial checkpointing with nsnap snapshots), which starts from them.  Torch only allocates and moves memory here; every
step and every adjoint step runs in libeuler2d_hip.  Ghost and corner entries
of the gradient are zero (ghosts follow the interior through the boundary
conditions; corners are treated as constants).
"""
from __future__ import annotations

import torch

from .hydro import PREC_FP64, HydroRun


def state_tensor(run: HydroRun, buf: int) -> torch.Tensor:
    """Copy of state buffer `buf` as a (4, jsize, pitch) float64 CUDA tensor."""
    if run.params.precision != PREC_FP64:
        raise ValueError("the autograd harness needs precision 0 (fp64)")
    _, pitch = run.device_buffer(buf)
    t = torch.empty((4, run.params.jsize, pitch), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    run.copy_from_buffer(buf, t.data_ptr())
    return t


def check_state_tensor(run: HydroRun, t: torch.Tensor, what: str):
    """Raise ValueError unless `t` can be handed to the library as a whole FP64
    state or field of `run`: a CUDA float64 tensor of shape (4, jsize, pitch)
    (the device layout), and the run in precision 0 (fp64).  The library copies
    state_bytes / field_bytes from such an address without knowing the tensor,
    so anything smaller would be read or written past its end."""
    if run.params.precision != PREC_FP64:
        raise ValueError("the autograd harness needs precision 0 (fp64), "
                         f"got {run.params.precision}")
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{what} must be a torch.Tensor, got {type(t).__name__}")
    if t.dtype != torch.float64:
        raise ValueError(f"{what} must be float64, got {t.dtype}")
    _, pitch = run.device_buffer(0)
    want = (4, run.params.jsize, pitch)
    if tuple(t.shape) != want:
        raise ValueError(f"{what} must have the device layout shape {want} (4, jsize, pitch), "
                         f"got {tuple(t.shape)}")
    if not t.is_cuda:
        raise ValueError(f"{what} must be a CUDA tensor, got device {t.device}")


def interior_mask(run: HydroRun, like: torch.Tensor) -> torch.Tensor:
    m = torch.zeros_like(like, dtype=torch.bool)
    m[:, 2:run.params.ny + 2, 2:run.params.nx + 2] = True
    return m


def record_dts(run: HydroRun, u0: torch.Tensor, nsteps: int) -> list[float]:
    """Forward run from u0 with the reference's dt rule (compute_dt each
    step, running maximum reset first); returns the dt sequence."""
    check_state_tensor(run, u0, "u0")
    run.invdt = 0.0
    torch.cuda.synchronize()
    run.copy_to_buffer(0, u0.data_ptr())
    dts = []
    for n in range(nsteps):
        dt = run.compute_dt(n % 2)
        dts.append(dt)
        run.godunov_unsplit(n, dt)
    run.synchronize()
    return dts


class GodunovRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0: torch.Tensor, run: HydroRun, dts, nsnap: int):
        check_state_tensor(run, u0, "u0")
        u0 = u0.contiguous()
        torch.cuda.synchronize()
        run.copy_to_buffer(0, u0.data_ptr())
        # the steps, keeping the snapshots the backward's reverse sweep starts from
        run.rollout_forward(0, [float(d) for d in dts], int(nsnap))
        out = torch.empty_like(u0)
        run.copy_from_buffer(len(dts) % 2, out.data_ptr())
        ctx.run, ctx.dts, ctx.nsnap = run, [float(d) for d in dts], int(nsnap)
        ctx.save_for_backward(u0)
        return out

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (u0,) = ctx.saved_tensors
        run = ctx.run
        check_state_tensor(run, grad_out, "grad_out")
        g = grad_out.contiguous()
        lam0 = torch.zeros_like(g)
        torch.cuda.synchronize()
        run.copy_to_buffer(0, u0.data_ptr())
        run.adjoint_rollout(0, ctx.dts, ctx.nsnap, g.data_ptr(), lam0.data_ptr())
        lam0 = torch.where(interior_mask(run, lam0), lam0, torch.zeros_like(lam0))
        return lam0, None, None, None
