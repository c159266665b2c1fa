"""Python binding of libeuler2d_hip (ctypes over the C-ABI in include/euler2d_hip.h).

Mirrors the reference's HydroParameters / HydroRun modules
(cuf/m_HydroParameters.cuf, cuf/m_HydroRun_gpu.cuf, serial_ref/m_run.f90):
`read_namelist` reads the reference's parameter file, `HydroParams` holds the
&MESH/&HYDRO/&OTHER values, and `HydroRun` owns the device state and offers
compute_dt / make_boundaries / godunov_unsplit with the reference's meaning.
All compute runs in the HIP library; nothing here falls back to the CPU.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
import pathlib
import re

import numpy as np

PKG = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG / "lib" / "libeuler2d_hip.so"

IMPL_FUSED, IMPL_MATERIALISED, IMPL_FLUX_STORE = 0, 1, 2
PREC_FP64, PREC_FP32, PREC_FP32_NATIVE, PREC_MIXED = 0, 1, 2, 3
RIEMANN_APPROX, RIEMANN_HLLC = 0, 2
PROBLEM_IMPLODE, PROBLEM_DIAG, PROBLEM_BUMPS = 0, 1, 2
STAGES = {"q": 0, "fx": 1, "fy": 2, "dqx": 3, "dqy": 4, "qxm": 5, "qxp": 6, "qym": 7, "qyp": 8}
(FN_PRIM, FN_SLOPE, FN_TRACE, FN_RIEMANN, FN_HLLC, FN_CMPFLX, FN_EOS, FN_DIV, FN_SQRT, FN_DIV_ROOT,
 FN_DIV_NEAR, FN_CONV_GT) = range(1, 13)
FN_WIDTH = {FN_PRIM: (4, 5), FN_SLOPE: (20, 8), FN_TRACE: (15, 4), FN_RIEMANN: (8, 8),
            FN_HLLC: (8, 8), FN_CMPFLX: (4, 4), FN_EOS: (2, 2), FN_DIV: (2, 1), FN_SQRT: (1, 1),
            FN_DIV_ROOT: (2, 2), FN_DIV_NEAR: (3, 1),
            FN_CONV_GT: (2, 1)}
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "fp32_native": PREC_FP32_NATIVE, "mixed": PREC_MIXED}
PROBLEMS = {"implode": PROBLEM_IMPLODE, "diag": PROBLEM_DIAG, "bumps": PROBLEM_BUMPS}


class LibraryNotBuilt(RuntimeError):
    pass


class E2DError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"euler2d_hip error {code}: {msg}")
        self.code = code


class _Params(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
                ("boundary_type_xmin", ctypes.c_int32), ("boundary_type_xmax", ctypes.c_int32),
                ("boundary_type_ymin", ctypes.c_int32), ("boundary_type_ymax", ctypes.c_int32),
                ("slope_type", ctypes.c_int32), ("niter_riemann", ctypes.c_int32),
                ("iorder", ctypes.c_int32), ("implementation_version", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("riemann_solver", ctypes.c_int32),
                ("gamma0", ctypes.c_double), ("cfl", ctypes.c_double),
                ("smallr", ctypes.c_double), ("smallc", ctypes.c_double)]


_lib = None

# every entry point of include/euler2d_hip.h with its ctypes signature
_SIGS = {
    "e2d_version": (ctypes.c_int, []),
    "e2d_device_available": (ctypes.c_int, []),
    "e2d_last_error": (ctypes.c_char_p, []),
    "e2d_create": (ctypes.c_int, [ctypes.POINTER(_Params), ctypes.POINTER(ctypes.c_void_p)]),
    "e2d_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "e2d_init_problem": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64]),
    "e2d_upload": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "e2d_download": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "e2d_compute_dt": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "e2d_make_boundaries": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "e2d_godunov_unsplit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]),
    "e2d_run_steps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "e2d_get_dt": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "e2d_get_invdt": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "e2d_set_invdt": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "e2d_sums": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "e2d_download_stage": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "e2d_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "e2d_enable_timers": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "e2d_get_timers": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "e2d_device_buffer": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
                                         ctypes.POINTER(ctypes.c_int64)]),
    "e2d_pointwise": (ctypes.c_int, [ctypes.POINTER(_Params), ctypes.c_int, ctypes.c_int,
                                     ctypes.c_void_p, ctypes.c_void_p]),
    "e2d_tangent_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                                        ctypes.c_void_p]),
    "e2d_adjoint_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                                        ctypes.c_void_p]),
    "e2d_adjoint_rollout": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                           ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.POINTER(ctypes.c_int64)]),
    "e2d_rollout_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                           ctypes.c_int]),
    "e2d_field_alloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "e2d_field_free": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "e2d_field_upload": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "e2d_field_download": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "e2d_copy_to_buffer": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "e2d_copy_from_buffer": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "e2d_run_steps_noalloc": (ctypes.c_int, [ctypes.POINTER(_Params), ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                             ctypes.c_void_p]),
}


def load_library(path: os.PathLike | None = None):
    """Load the in-tree HIP library; raise LibraryNotBuilt when it is absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = pathlib.Path(path) if path else LIB_PATH
    if not p.exists():
        raise LibraryNotBuilt(f"{p} is not built: run `make -C euler2d_cudafortran_amd` "
                              "or __graft_entry__.build()")
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return list(_SIGS)


def _check(rc: int):
    if rc != 0:
        raise E2DError(rc, load_library().e2d_last_error().decode(errors="replace"))


def device_available() -> bool:
    return bool(load_library().e2d_device_available())


# ---------------------------------------------------------------------------
# parameter file (Fortran namelist, serial_ref/test.nml / cuf/test.nml layout)

def _parse_value(tok: str):
    t = tok.strip().rstrip(",").strip()
    if not t:
        return None
    if t[0] in "'\"":
        return t[1:-1]
    tl = t.lower()
    if tl in (".true.", "t", ".t."):
        return True
    if tl in (".false.", "f", ".f."):
        return False
    try:
        return int(t)
    except ValueError:
        return float(tl.replace("d", "e"))


_PAIR_RE = re.compile(r"(\w+)\s*=\s*('[^']*'|\"[^\"]*\"|[^,\s/]+)")


def read_namelist(path: os.PathLike) -> dict:
    """{group: {name: value}} with lower-case keys, for the subset of namelist
    syntax the reference's parameter files use (one group per '&name ... /'
    block, scalar name=value pairs, '!' comments)."""
    out: dict = {}
    group = None
    for raw in pathlib.Path(path).read_text().splitlines():
        line = raw.split("!")[0].strip()
        if not line:
            continue
        if line.startswith("&"):
            name = line[1:].split()[0].lower() if len(line) > 1 else ""
            if name in ("end", ""):
                group = None
                continue
            group = out.setdefault(name, {})
            line = line[1 + len(name):]
        if group is not None:
            for k, v in _PAIR_RE.findall(line):
                group[k.lower()] = _parse_value(v)
        if line.endswith("/") or line == "/":
            group = None
    return out


@dataclasses.dataclass
class HydroParams:
    """&MESH, &HYDRO and &OTHER values (serial_ref/m_parameters.f90:10-36,
    cuf/m_HydroParameters.cuf:17-80); loop controls (&RUN) kept beside them."""
    nx: int = 128
    ny: int = 128
    boundary_type_xmin: int = 1
    boundary_type_xmax: int = 1
    boundary_type_ymin: int = 1
    boundary_type_ymax: int = 1
    gamma0: float = 1.4
    cfl: float = 0.8
    smallr: float = 1e-7
    smallc: float = 1e-7
    niter_riemann: int = 10
    iorder: int = 2
    slope_type: int = 1
    implementation_version: int = IMPL_FUSED
    precision: int = PREC_FP64
    riemann_solver: int = RIEMANN_APPROX
    problem: str = "implode"
    tEnd: float = 10.0
    nStepmax: int = 100
    nOutput: int = 1000

    @property
    def isize(self) -> int:
        return self.nx + 4

    @property
    def jsize(self) -> int:
        return self.ny + 4

    @classmethod
    def from_namelist(cls, path: os.PathLike) -> "HydroParams":
        nml = read_namelist(path)
        p = cls()
        run, mesh, hydro, other = (nml.get(g, {}) for g in ("run", "mesh", "hydro", "other"))
        for k, attr in (("tend", "tEnd"), ("nstepmax", "nStepmax"), ("noutput", "nOutput")):
            if k in run:
                setattr(p, attr, run[k])
        for k in ("nx", "ny", "boundary_type_xmin", "boundary_type_xmax", "boundary_type_ymin",
                  "boundary_type_ymax"):
            if k in mesh:
                setattr(p, k, int(mesh[k]))
        for k in ("gamma0", "cfl", "smallr", "smallc"):
            if k in hydro:
                setattr(p, k, float(hydro[k]))
        for k in ("niter_riemann", "iorder", "slope_type"):
            if k in hydro:
                setattr(p, k, int(hydro[k]))
        if "implementationversion" in other:
            p.implementation_version = int(other["implementationversion"])
        if "precision" in other:
            v = other["precision"]
            p.precision = PRECISIONS[v.lower()] if isinstance(v, str) else int(v)
        if "riemann_solver" in other:
            p.riemann_solver = int(other["riemann_solver"])
        if "problem" in other:
            p.problem = str(other["problem"]).lower()
        return p

    def to_c(self) -> _Params:
        return _Params(self.nx, self.ny, self.boundary_type_xmin, self.boundary_type_xmax,
                       self.boundary_type_ymin, self.boundary_type_ymax, self.slope_type,
                       self.niter_riemann, self.iorder, self.implementation_version,
                       self.precision, self.riemann_solver, self.gamma0, self.cfl, self.smallr,
                       self.smallc)


class HydroRun:
    """Device state of one run: buffers u (0) and u2 (1) in the reference's
    Fortran layout (isize, jsize, 4), the running invDt and the step counter."""

    def __init__(self, params: HydroParams):
        self.params = params
        self.lib = load_library()
        self._cp = params.to_c()
        h = ctypes.c_void_p()
        _check(self.lib.e2d_create(ctypes.byref(self._cp), ctypes.byref(h)))
        self.ctx = h
        self.nStep = 0
        self._fields: set[int] = set()  # device fields from field_alloc, freed by close()

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "ctx", None):
            for d in getattr(self, "_fields", ()):
                self.lib.e2d_field_free(self.ctx, d)
            self._fields = set()
            self.lib.e2d_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        return (self.params.isize, self.params.jsize, 4)

    # -- reference procedures ----------------------------------------------
    def init(self, problem: str | int | None = None, seed: int = 0):
        """initHydroRun's initial condition (on the device) + step counter reset."""
        pr = self.params.problem if problem is None else problem
        code = PROBLEMS[pr] if isinstance(pr, str) else int(pr)
        _check(self.lib.e2d_init_problem(self.ctx, code, seed))
        self.nStep = 0

    def upload(self, buf: int, u: np.ndarray):
        a = np.asfortranarray(u, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"expected shape {self.shape}, got {a.shape}")
        flat = np.ascontiguousarray(a.reshape(-1, order="F"))
        _check(self.lib.e2d_upload(self.ctx, buf, flat.ctypes.data))

    def download(self, buf: int) -> np.ndarray:
        out = np.empty(self.shape[0] * self.shape[1] * 4, dtype=np.float64)
        _check(self.lib.e2d_download(self.ctx, buf, out.ctypes.data))
        return out.reshape(self.shape, order="F")

    def download_stage(self, name: str) -> np.ndarray:
        out = np.empty(self.shape[0] * self.shape[1] * 4, dtype=np.float64)
        _check(self.lib.e2d_download_stage(self.ctx, STAGES[name], out.ctypes.data))
        return out.reshape(self.shape, order="F")

    def compute_dt(self, use_u: int) -> float:
        dt = ctypes.c_double()
        _check(self.lib.e2d_compute_dt(self.ctx, use_u, ctypes.byref(dt)))
        return dt.value

    def make_boundaries(self, buf: int):
        _check(self.lib.e2d_make_boundaries(self.ctx, buf))

    def godunov_unsplit(self, n_step: int, dt: float):
        _check(self.lib.e2d_godunov_unsplit(self.ctx, n_step, dt))

    def run_steps(self, nsteps: int) -> int:
        n = ctypes.c_int(self.nStep)
        _check(self.lib.e2d_run_steps(self.ctx, nsteps, ctypes.byref(n)))
        self.nStep = n.value
        return self.nStep

    def get_dt(self) -> float:
        dt = ctypes.c_double()
        _check(self.lib.e2d_get_dt(self.ctx, ctypes.byref(dt)))
        return dt.value

    @property
    def invdt(self) -> float:
        v = ctypes.c_double()
        _check(self.lib.e2d_get_invdt(self.ctx, ctypes.byref(v)))
        return v.value

    @invdt.setter
    def invdt(self, v: float):
        _check(self.lib.e2d_set_invdt(self.ctx, v))

    def sums(self, buf: int) -> np.ndarray:
        out = np.zeros(4, dtype=np.float64)
        _check(self.lib.e2d_sums(self.ctx, buf, out.ctypes.data))
        return out

    def synchronize(self):
        _check(self.lib.e2d_synchronize(self.ctx))

    def enable_timers(self, on: bool = True):
        _check(self.lib.e2d_enable_timers(self.ctx, 1 if on else 0))

    def get_timers(self) -> dict:
        """{'godunov'|'boundaries'|'dt': (milliseconds, launches)} since the last call."""
        ms = np.zeros(3, dtype=np.float64)
        cnt = np.zeros(3, dtype=np.int64)
        _check(self.lib.e2d_get_timers(self.ctx, ms.ctypes.data, cnt.ctypes.data))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(("godunov", "boundaries", "dt"))}

    def device_buffer(self, buf: int):
        p = ctypes.c_void_p()
        pitch = ctypes.c_int64()
        _check(self.lib.e2d_device_buffer(self.ctx, buf, ctypes.byref(p), ctypes.byref(pitch)))
        return p.value, pitch.value

    # -- tangent (FP64) / adjoint (all precisions), two-shock solver --------
    def field_alloc(self) -> int:
        """Zero FP64 vector field in the device layout; returns its device address
        (freed by field_free or when the run is closed)."""
        p = ctypes.c_void_p()
        _check(self.lib.e2d_field_alloc(self.ctx, ctypes.byref(p)))
        self._fields.add(p.value)
        return p.value

    def field_free(self, d: int):
        _check(self.lib.e2d_field_free(self.ctx, d))
        self._fields.discard(d)

    def field_upload(self, d: int, h: np.ndarray):
        a = np.ascontiguousarray(np.asfortranarray(h, dtype=np.float64).reshape(-1, order="F"))
        _check(self.lib.e2d_field_upload(self.ctx, d, a.ctypes.data))

    def field_download(self, d: int) -> np.ndarray:
        out = np.empty(self.shape[0] * self.shape[1] * 4, dtype=np.float64)
        _check(self.lib.e2d_field_download(self.ctx, d, out.ctypes.data))
        return out.reshape(self.shape, order="F")

    def tangent_step(self, buf: int, dt: float, d_v_in: int, d_v_out: int):
        _check(self.lib.e2d_tangent_step(self.ctx, buf, dt, d_v_in, d_v_out))

    def adjoint_step(self, buf: int, dt: float, d_lam_in: int, d_lam_out: int):
        _check(self.lib.e2d_adjoint_step(self.ctx, buf, dt, d_lam_in, d_lam_out))

    def adjoint_rollout(self, n_step0: int, dts, nsnap: int, d_lam_n: int, d_lam_0: int) -> int:
        """lam_0 = (dU_n/dU_0)^T lam_n over len(dts) frozen-dt steps from the state in
        buffer n_step0 % 2; returns the number of forward steps recomputed."""
        d = np.ascontiguousarray(dts, dtype=np.float64)
        nf = ctypes.c_int64()
        _check(self.lib.e2d_adjoint_rollout(self.ctx, n_step0, d.size, d.ctypes.data, nsnap, d_lam_n, d_lam_0,
                                            ctypes.byref(nf)))
        return nf.value

    def rollout_forward(self, n_step0: int, dts, nsnap: int):
        """len(dts) frozen-dt steps from the state in buffer n_step0 % 2, keeping in
        the nsnap snapshots the states that the next adjoint_rollout with the same
        arguments (and the same starting state back in its buffer) starts from."""
        d = np.ascontiguousarray(dts, dtype=np.float64)
        _check(self.lib.e2d_rollout_forward(self.ctx, n_step0, d.size, d.ctypes.data, nsnap))

    def copy_to_buffer(self, buf: int, d_src: int):
        _check(self.lib.e2d_copy_to_buffer(self.ctx, buf, d_src))

    def copy_from_buffer(self, buf: int, d_dst: int):
        _check(self.lib.e2d_copy_from_buffer(self.ctx, buf, d_dst))

    def current(self) -> int:
        """Buffer holding the latest state (u for even nStep, u2 for odd)."""
        return self.nStep % 2

    # -- main loop of the reference (serial_ref/main.f90:42-76) ---------------
    def run_reference_loop(self, nstepmax: int, tend: float = 10.0) -> list[float]:
        """Host-driven loop: compute_dt each step (synchronous) then godunov_unsplit.
        t is never advanced, as in the reference's loop (serial_ref/main.f90:53-76),
        so nstepmax alone ends it."""
        dts = []
        t = 0.0
        dt = self.compute_dt(self.nStep % 2)
        self.make_boundaries(0)
        while t < tend and self.nStep < nstepmax:
            dt = self.compute_dt(self.nStep % 2)
            dts.append(dt)
            self.godunov_unsplit(self.nStep, dt)
            self.nStep += 1
        return dts


def run_steps_noalloc(params: HydroParams, d_u: int, d_u2: int, pitch: int, nsteps: int, n_step: int,
                      d_invdt: int, stream: int = 0):
    """e2d_run_steps_noalloc: nsteps device-resident steps on caller-owned device
    buffers (addresses d_u, d_u2 in the device layout, row pitch `pitch`), the
    running invDt in the device double at d_invdt; enqueued on `stream`
    (0 = the null stream), nothing allocated and nothing waited for."""
    cp = params.to_c()
    _check(load_library().e2d_run_steps_noalloc(ctypes.byref(cp), d_u, d_u2, pitch, nsteps, n_step, d_invdt,
                                                stream))


def pointwise(params: HydroParams, func: int, x: np.ndarray) -> np.ndarray:
    """Device pointwise numerics over a batch (known-answer tests)."""
    nin, nout = FN_WIDTH[func]
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, nin)
    y = np.zeros((x.shape[0], nout), dtype=np.float64)
    cp = params.to_c()
    _check(load_library().e2d_pointwise(ctypes.byref(cp), func, x.shape[0], x.ctypes.data,
                                        y.ctypes.data))
    return y
