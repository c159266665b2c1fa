import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def golden():
    return GOLDEN


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library with a usable device; GPU tests fail (not skip) without it."""
    import euler2d_cudafortran_amd as e2d
    lib = e2d.load_library()
    assert e2d.device_available(), "no HIP device visible to libeuler2d_hip"
    return lib
