"""libeuler2d_hip on the CPU: the library loads, exports every entry point
include/euler2d_hip.h declares, and rejects bad arguments before any HIP call
(the reference stops on a bad implementationVersion,
cuf/m_HydroParameters.cuf:128-135; here the C-ABI returns E2D_ERR_ARG with the
message).  No compute call is made here."""
import ctypes
import re

import pytest

import euler2d_cudafortran_amd as e2d
from conftest import ROOT

HEADER = ROOT / "include" / "euler2d_hip.h"


def _declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(e2d_\w+)\s*\(", text)))


def test_header_and_binding_agree():
    declared = _declared()
    assert len(declared) >= 20
    assert sorted(e2d.exported_symbols()) == declared


def test_library_exports_every_symbol():
    lib = e2d.load_library()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.e2d_version() == 1


def _params(**kw):
    p = e2d.HydroParams(nx=16, ny=16)
    for k, v in kw.items():
        setattr(p, k, v)
    return p.to_c()


@pytest.mark.parametrize("bad, msg", [
    (dict(implementation_version=3), "implementation version"),
    (dict(implementation_version=7), "implementation version"),
    (dict(nx=0), "nx and ny"),
    (dict(ny=3), "nx and ny"),
    (dict(precision=5), "precision"),
    (dict(riemann_solver=1), "riemann_solver"),
    (dict(gamma0=1.0), "gamma0"),
    (dict(cfl=0.0), "cfl"),
    (dict(smallr=0.0), "smallr"),
    (dict(smallc=1e-150), "floors below the double range"),
    (dict(smallr=1e-100, smallc=1e-80), "floors below the double range"),
    (dict(gamma0=1.0 + 1e-300), "gamma0"),
    (dict(smallr=1e10, smallc=1e-150), "Newton stop test"),
    (dict(niter_riemann=-1), "niter_riemann"),
])
def test_create_rejects_bad_params(bad, msg):
    lib = e2d.load_library()
    cp = _params(**bad)
    h = ctypes.c_void_p()
    rc = lib.e2d_create(ctypes.byref(cp), ctypes.byref(h))
    assert rc == -1
    assert h.value is None
    assert msg in lib.e2d_last_error().decode()


def test_null_arguments_rejected():
    lib = e2d.load_library()
    assert lib.e2d_create(None, None) == -1
    h = ctypes.c_void_p()
    assert lib.e2d_create(None, ctypes.byref(h)) == -1
    assert "NULL" in lib.e2d_last_error().decode()
    d = ctypes.c_double()
    for rc in (lib.e2d_compute_dt(None, 0, ctypes.byref(d)), lib.e2d_make_boundaries(None, 0),
               lib.e2d_godunov_unsplit(None, 0, 1e-3), lib.e2d_sums(None, 0, None),
               lib.e2d_get_dt(None, ctypes.byref(d)), lib.e2d_synchronize(None)):
        assert rc == -1
    assert lib.e2d_destroy(None) == 0
    cp = _params()
    assert lib.e2d_pointwise(ctypes.byref(cp), 99, 1, None, None) == -1
    assert lib.e2d_pointwise(ctypes.byref(cp), 1, 0, None, None) == 0
    # the function codes end at the last one the header declares
    last = max(e2d.hydro.FN_WIDTH)
    assert lib.e2d_pointwise(ctypes.byref(cp), last + 1, 1, None, None) == -1
    assert "unknown function" in lib.e2d_last_error().decode()
    assert lib.e2d_pointwise(ctypes.byref(cp), 0, 1, None, None) == -1
    assert lib.e2d_pointwise(ctypes.byref(cp), last, 0, None, None) == 0


def test_namelist_reference_layout(tmp_path):
    """The reference parameter file (cuf/test.nml layout) plus the &OTHER
    extensions of this implementation."""
    f = tmp_path / "test.nml"
    f.write_text("&RUN\ntEnd=10.0\nnStepmax=5000\nnOutput=1000\n/\n\n"
                 "&MESH\nnx=1024\nny=512\nboundary_type_xmin=1\nboundary_type_xmax=2\n"
                 "boundary_type_ymin=3\nboundary_type_ymax=1\n/\n\n"
                 "&HYDRO\ngamma0=1.666\ncfl=0.8\nsmallr=1e-7\nsmallc=1e-7\nniter_riemann=10\n"
                 "iorder=2\nslope_type=2\n/\n\n"
                 "&OTHER\nimplementationVersion=1\nprecision='fp32'\nproblem='diag'  ! comment\n/\n")
    p = e2d.HydroParams.from_namelist(f)
    assert (p.nx, p.ny, p.nStepmax, p.nOutput) == (1024, 512, 5000, 1000)
    assert (p.boundary_type_xmax, p.boundary_type_ymin) == (2, 3)
    assert p.gamma0 == 1.666 and p.slope_type == 2 and p.smallr == 1e-7
    assert p.implementation_version == 1 and p.precision == e2d.PREC_FP32 and p.problem == "diag"


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(e2d.LibraryNotBuilt):
        e2d.load_library(tmp_path / "nope.so")


def test_trace_ranges_linked():
    """The library opens roctx ranges (recorded by rocprofv3 --marker-trace) named
    after the reference's NVTX ranges (serial_ref/m_run.f90:195,264,336,387,502,571)."""
    import subprocess
    so = e2d.LIB_PATH
    undef = subprocess.run(["nm", "-D", "--undefined-only", str(so)], capture_output=True, text=True,
                           check=True).stdout
    assert "roctxRangePushA" in undef and "roctxRangePop" in undef
    data = so.read_bytes()
    for name in (b"godunov_compute", b"compute_dt", b"convertToPrimitives", b"computeAndStoreFluxes",
                 b"updateHydro", b"make_boundaries"):
        assert name + b"\0" in data, name


def test_binding_constants_match_the_header():
    """Every E2D_IMPL_/PREC_/RIEMANN_/PROBLEM_/STAGE_/FN_ value of
    include/euler2d_hip.h has the same value in the Python binding, and
    each pointwise function has its widths there."""
    import euler2d_cudafortran_amd.hydro as hy

    text = (ROOT / "include" / "euler2d_hip.h").read_text()
    defs = {m.group(1): int(m.group(2)) for m in re.finditer(r"^#define E2D_(\w+) \(?(-?\d+)\)?", text, re.M)}
    checked = 0
    for name, value in defs.items():
        group = name.split("_")[0]
        if group in ("IMPL", "PREC", "RIEMANN", "PROBLEM", "FN"):
            assert getattr(hy, name) == value, name
            checked += 1
        elif group == "STAGE":
            assert hy.STAGES[name[len("STAGE_"):].lower()] == value, name
            checked += 1
    fns = {v for k, v in defs.items() if k.startswith("FN_")}
    assert fns == set(hy.FN_WIDTH) and checked >= 30
