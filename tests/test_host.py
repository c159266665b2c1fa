"""The Fortran host program (euler2d_cudafortran_amd/bin/euler2d_hip).

CPU: argument and parameter errors stop before any GPU work (the reference's
usage message and implementationVersion check, serial_ref/m_run.f90:76-82,
cuf/m_HydroParameters.cuf:128-135; the library's validation message).
GPU: the program's VTK files are byte-identical to the files the reference
program itself writes (serial_ref/main.f90 compiled with the same Fortran
compiler; tests/golden/vtk_*.vti), and a 128^2 run reproduces the reference's
final state bit for bit.
"""
import pathlib
import struct
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from oracle_lib import write_nml

EXE = ROOT / "euler2d_cudafortran_amd" / "bin" / "euler2d_hip"


def _run(args, cwd, timeout=120):
    return subprocess.run([str(EXE), *args], cwd=cwd, capture_output=True, text=True, timeout=timeout)


@pytest.fixture(scope="module")
def exe():
    assert EXE.exists(), f"{EXE} not built (make -C euler2d_cudafortran_amd host)"
    return EXE


def test_usage_without_argument(exe, tmp_path):
    r = _run([], tmp_path)
    assert "You should provide an input parameter file" in r.stdout
    assert "Usage:" in r.stdout


@pytest.mark.parametrize("version", [3, 7, -1])
def test_bad_implementation_version(exe, tmp_path, version):
    write_nml(tmp_path / "t.nml", 16, extra=f"&OTHER\nimplementationVersion={version}\n/\n")
    r = _run(["t.nml"], tmp_path)
    # the reference's message, word for word (cuf/m_HydroParameters.cuf:130)
    assert "The implementation version parameter should 0, 1 or 2 !!!" in r.stdout
    assert not list(tmp_path.glob("*.vti"))


def test_bad_mesh_reports_library_message(exe, tmp_path):
    write_nml(tmp_path / "t.nml", 2)
    r = _run(["t.nml"], tmp_path)
    assert r.returncode != 0
    assert "e2d_create failed" in r.stdout and "nx and ny must be >= 4" in r.stdout


def test_missing_parameter_file(exe, tmp_path):
    r = _run(["nope.nml"], tmp_path)
    assert r.returncode != 0
    assert "Cannot open parameter file" in r.stdout


# ---------------------------------------------------------------------------
# GPU

def _read_vti_ascii(path: pathlib.Path, nx: int, ny: int) -> np.ndarray:
    """(nx, ny, 4) cell data of an ASCII .vti written by saveVTK."""
    text = path.read_text()
    out = np.zeros((nx, ny, 4))
    parts = text.split("<DataArray")[1:]
    assert len(parts) == 4
    for v, part in enumerate(parts):
        body = part.split(">", 1)[1].split("</DataArray>")[0]
        vals = np.array([float(t) for t in body.split()])
        assert vals.size == nx * ny
        out[:, :, v] = vals.reshape((ny, nx)).T
    return out


def _read_vti_binary(path: pathlib.Path, nx: int, ny: int, dtype) -> np.ndarray:
    data = path.read_bytes()
    start = data.index(b"_", data.index(b"<AppendedData")) + 1
    out = np.zeros((nx, ny, 4))
    off = start
    for v in range(4):
        (nbytes,) = struct.unpack_from("<Q", data, off)
        off += 8
        arr = np.frombuffer(data, dtype=dtype, count=nbytes // np.dtype(dtype).itemsize, offset=off)
        off += nbytes
        out[:, :, v] = arr.reshape((ny, nx)).T
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["fp64", "fp32"])
@pytest.mark.parametrize("impl", [0, 1, 2])
def test_vtk_files_byte_identical_to_reference_program(gpu_lib, exe, tmp_path, kind, impl):
    write_nml(tmp_path / "test.nml", 16, 16, nstepmax=3, noutput=2, gamma0=1.4, slope_type=1,
              extra=f"&OTHER\nimplementationVersion={impl}\nprecision='{kind}'\n/\n")
    r = _run(["test.nml"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    for step in ("00000000", "00000002", "00000003"):
        mine = (tmp_path / f"euler2d_{step}.vti").read_bytes()
        ref = (GOLDEN / f"vtk_{kind}_{step}.vti").read_bytes()
        assert mine == ref, f"euler2d_{step}.vti differs from the reference program's"


@pytest.mark.gpu
@pytest.mark.parametrize("device_loop", [".false.", ".true."])
def test_host_128_diag_matches_reference_state(gpu_lib, exe, tmp_path, device_loop):
    d = np.load(GOLDEN / "run_diag128_fp64.npz")
    write_nml(tmp_path / "t.nml", 128, nstepmax=200, noutput=1000,
              extra=f"&OTHER\nproblem='diag'\ndevice_loop={device_loop}\n/\n")
    r = _run(["t.nml"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Perf" in r.stdout
    final = _read_vti_ascii(tmp_path / "euler2d_00000200.vti", 128, 128)
    assert np.array_equal(final, d["u_end"][2:130, 2:130, :])


@pytest.mark.gpu
@pytest.mark.parametrize("kind,dtype", [("fp64", "<f8"), ("fp32", "<f4")])
def test_binary_vtk_matches_ascii(gpu_lib, exe, tmp_path, kind, dtype):
    for fmt in ("ascii", "binary"):
        wd = tmp_path / fmt
        wd.mkdir()
        write_nml(wd / "t.nml", 40, 24, nstepmax=7, noutput=5,
                  extra=f"&OTHER\nproblem='diag'\nprecision='{kind}'\nvtk_format='{fmt}'\n/\n")
        r = _run(["t.nml"], wd)
        assert r.returncode == 0, r.stdout + r.stderr
    for step in ("00000000", "00000005", "00000007"):
        a = _read_vti_ascii(tmp_path / "ascii" / f"euler2d_{step}.vti", 40, 24)
        b = _read_vti_binary(tmp_path / "binary" / f"euler2d_{step}.vti", 40, 24, dtype)
        if kind == "fp32":  # the ASCII file holds the shortest decimal of each float
            a, b = a.astype(np.float32), b.astype(np.float32)
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_async_output_writes_the_same_files(gpu_lib, exe, tmp_path):
    """async_output (with device_loop) launches the steps up to the next output
    before writing the current file: the files are byte-identical to the
    serial order's, and the Perf line is still printed."""
    files = {}
    for mode in (".false.", ".true."):
        wd = tmp_path / mode.strip(".")
        wd.mkdir()
        write_nml(wd / "t.nml", 64, 48, nstepmax=23, noutput=5,
                  extra=f"&OTHER\nproblem='diag'\ndevice_loop=.true.\nasync_output={mode}\n/\n")
        r = _run(["t.nml"], wd)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "Perf" in r.stdout
        files[mode] = {p.name: p.read_bytes() for p in sorted(wd.glob("*.vti"))}
    assert list(files[".true."]) == [f"euler2d_{s:08d}.vti" for s in (0, 5, 10, 15, 20, 23)]
    assert files[".true."] == files[".false."]


@pytest.mark.gpu
def test_adjoint_from_fortran_matches_python(gpu_lib, tmp_path):
    """The discrete adjoint called from Fortran (bin/e2d_adjoint_check, through
    the bind(C) interfaces of module euler2d_hip: recorded forward, then the
    checkpointed reverse sweep) gives the Python API's gradient and recompute
    count for the same rollout."""
    import euler2d_cudafortran_amd as e2d

    adj = ROOT / "euler2d_cudafortran_amd" / "bin" / "e2d_adjoint_check"
    if not adj.exists():
        pytest.skip("Fortran adjoint check not built (no Fortran compiler)")
    nx, ny, nsteps, nsnap = 64, 48, 20, 6
    r = subprocess.run([str(adj), str(nx), str(ny), str(nsteps), str(nsnap)], capture_output=True, text=True,
                       timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = dict(line.split() for line in r.stdout.strip().splitlines())
    with e2d.HydroRun(e2d.HydroParams(nx=nx, ny=ny)) as run:
        run.init("diag")
        run.make_boundaries(0)
        dts = []
        for n in range(nsteps):
            dts.append(run.compute_dt(n % 2))
            run.godunov_unsplit(n, dts[-1])
        run.init("diag")
        run.make_boundaries(0)
        run.rollout_forward(0, dts, nsnap)
        lam = np.zeros((nx + 4, ny + 4, 4))
        lam[2:nx + 2, 2:ny + 2, 0] = 1.0
        d_n, d_0 = run.field_alloc(), run.field_alloc()
        run.field_upload(d_n, lam)
        run.init("diag")
        run.make_boundaries(0)
        nf = run.adjoint_rollout(0, dts, nsnap, d_n, d_0)
        g = run.field_download(d_0)
    assert int(lines["forward_steps_recomputed"]) == nf
    for v in range(4):
        want = float(np.sum(g[2:nx + 2, 2:ny + 2, v]))
        got = float(lines[f"gradient_sum_var{v + 1}"])
        assert abs(got - want) <= 1e-12 * max(1.0, abs(want)), (v, got, want)
