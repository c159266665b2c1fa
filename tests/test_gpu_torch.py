"""The torch.autograd harness on the GPU, in a fresh interpreter
(tests/helpers/torch_autograd_check.py explains why)."""
import subprocess
import sys

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_autograd_rollout(gpu_lib):
    r = subprocess.run([sys.executable, str(ROOT / "tests" / "helpers" / "torch_autograd_check.py")],
                       capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert "AUTOGRAD_CHECK ok" in r.stdout
