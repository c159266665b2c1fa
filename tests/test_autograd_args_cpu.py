"""The torch.autograd harness rejects tensors the library cannot take (wrong
precision, dtype, shape or device) before any library call; in a fresh
interpreter, tests/helpers/torch_autograd_args.py."""
import subprocess
import sys

from conftest import ROOT


def test_autograd_argument_checks():
    r = subprocess.run([sys.executable, str(ROOT / "tests" / "helpers" / "torch_autograd_args.py")],
                       capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert "AUTOGRAD_ARGS ok" in r.stdout
