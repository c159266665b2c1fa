"""Argument checks of the torch.autograd harness, on the CPU and in a fresh
interpreter (torch loads its HIP runtime before libeuler2d_hip can, see
torch_autograd_check.py).  No device is needed: the run is a stand-in that
answers the two questions the checks ask (precision, row pitch), and every
rejected call raises before any library call.  Prints "AUTOGRAD_ARGS ok"."""
import pathlib
import sys

import torch  # noqa: I001

sys.path.insert(0, str(pathlib.Path(__file__).resolve().parents[2]))
import euler2d_cudafortran_amd as e2d  # noqa: E402
from euler2d_cudafortran_amd.autograd import GodunovRollout, check_state_tensor, record_dts  # noqa: E402


class FakeRun:
    """Precision and row pitch of a run; any other attribute is a library call, which must not happen."""

    def __init__(self, precision, nx=8, pitch=64):
        self.params = e2d.HydroParams(nx=nx, ny=nx, precision=precision)
        self._pitch = pitch

    def device_buffer(self, buf):
        return 0, self._pitch


def expect(fn, msg):
    try:
        fn()
    except ValueError as e:
        assert msg in str(e), (msg, str(e))
        return
    raise AssertionError(f"no ValueError ({msg})")


def main():
    run = FakeRun(e2d.PREC_FP64)
    shape = (4, 12, 64)  # (4, jsize, pitch) of an 8x8 run with a 64-element pitch
    good = torch.zeros(shape, dtype=torch.float64)
    expect(lambda: check_state_tensor(FakeRun(e2d.PREC_FP32), good, "u0"), "precision 0")
    expect(lambda: check_state_tensor(FakeRun(e2d.PREC_MIXED), good, "u0"), "precision 0")
    expect(lambda: check_state_tensor(run, good.numpy(), "u0"), "torch.Tensor")
    expect(lambda: check_state_tensor(run, good.float(), "u0"), "float64")
    # the natural (4, jsize, isize) shape: the library would read past it
    expect(lambda: check_state_tensor(run, torch.zeros((4, 12, 12), dtype=torch.float64), "u0"), "shape")
    expect(lambda: check_state_tensor(run, torch.zeros((4, 12 * 64), dtype=torch.float64), "u0"), "shape")
    expect(lambda: check_state_tensor(run, good, "u0"), "CUDA")
    expect(lambda: GodunovRollout.apply(good, run, [1e-3], 1), "CUDA")
    expect(lambda: GodunovRollout.apply(good.float(), run, [1e-3], 1), "float64")
    expect(lambda: record_dts(run, torch.zeros((4, 12, 12), dtype=torch.float64), 1), "shape")
    if torch.cuda.is_available():
        check_state_tensor(run, good.cuda(), "u0")
    print("AUTOGRAD_ARGS ok")


if __name__ == "__main__":
    main()
