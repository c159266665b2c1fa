"""Test-side access to the oracle (oracle/liboracle.so) and to reference dumps.

Test infrastructure only: the oracle is the checker, never the thing measured
or shipped.  `Oracle` wraps the CPU restatement of the reference serial solver
(oracle/cpu_ref.h); `read_ref_dump` parses the record stream written by
oracle/ref_driver.f90 (the compiled reference); `write_nml` writes a namelist
file in the reference's format (serial_ref/test.nml).
"""
from __future__ import annotations

import ctypes
import os
import pathlib
import struct

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "liboracle.so"

U, U2, Q, FX, FY = 0, 1, 2, 3, 4
IC_IMPLODE, IC_DIAG = 0, 1
KAT_PRIM, KAT_SLOPE, KAT_TRACE, KAT_RIEMANN, KAT_HLLC, KAT_CMPFLX, KAT_EOS, KAT_ITERS = range(1, 9)
KAT_SHAPES = {KAT_PRIM: (4, 5), KAT_SLOPE: (20, 8), KAT_TRACE: (15, 4), KAT_RIEMANN: (8, 8),
              KAT_HLLC: (8, 8), KAT_CMPFLX: (4, 4), KAT_EOS: (2, 2), KAT_ITERS: (8, 1)}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            raise FileNotFoundError(f"{ORACLE_SO} missing: run `make -C oracle`")
        L = ctypes.CDLL(str(ORACLE_SO))
        dp = ctypes.POINTER(ctypes.c_double)
        for s in ("f64", "f32", "f80", "f128"):
            f = getattr(L, f"oracle_create_{s}")
            f.restype = ctypes.c_void_p
            f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                          ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                          ctypes.c_double, ctypes.c_double]
            getattr(L, f"oracle_destroy_{s}").argtypes = [ctypes.c_void_p]
            getattr(L, f"oracle_init_{s}").argtypes = [ctypes.c_void_p, ctypes.c_int]
            f = getattr(L, f"oracle_compute_dt_{s}")
            f.restype = ctypes.c_double
            f.argtypes = [ctypes.c_void_p, ctypes.c_int]
            f = getattr(L, f"oracle_get_invdt_{s}")
            f.restype = ctypes.c_double
            f.argtypes = [ctypes.c_void_p]
            getattr(L, f"oracle_set_invdt_{s}").argtypes = [ctypes.c_void_p, ctypes.c_double]
            getattr(L, f"oracle_make_boundaries_{s}").argtypes = [ctypes.c_void_p, ctypes.c_int]
            getattr(L, f"oracle_godunov_{s}").argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
            f = getattr(L, f"oracle_run_{s}")
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, dp]
            getattr(L, f"oracle_get_{s}").argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
            getattr(L, f"oracle_set_{s}").argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
            getattr(L, f"oracle_kat_{s}").argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, dp, dp]
            f = getattr(L, f"oracle_fd_{s}")
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_void_p, dp, dp, ctypes.c_int, dp, ctypes.c_double, dp, ctypes.c_int, dp]
        L.oracle_threads.restype = ctypes.c_int
        L.oracle_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """One run of the CPU restatement (u, u2, q, fx, fy and the running invDt)."""

    def __init__(self, nx, ny=None, bc=(1, 1, 1, 1), slope_type=1, niter_riemann=10,
                 riemann_solver=0, gamma0=1.4, cfl=0.8, smallr=1e-7, smallc=1e-7, precision="f64"):
        ny = nx if ny is None else ny
        self.s = precision
        self.nx, self.ny = nx, ny
        self.isize, self.jsize = nx + 4, ny + 4
        L = lib()
        bca = (ctypes.c_int * 4)(*bc)
        self.h = getattr(L, f"oracle_create_{self.s}")(nx, ny, bca, slope_type, niter_riemann,
                                                       riemann_solver, gamma0, cfl, smallr, smallc)
        if not self.h:
            raise MemoryError("oracle_create failed")

    def _f(self, name):
        return getattr(lib(), f"oracle_{name}_{self.s}")

    def close(self):
        if self.h:
            self._f("destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self, problem=IC_DIAG):
        self._f("init")(self.h, problem)

    def compute_dt(self, use_u: int) -> float:
        return self._f("compute_dt")(self.h, use_u)

    @property
    def invdt(self) -> float:
        return self._f("get_invdt")(self.h)

    @invdt.setter
    def invdt(self, v: float):
        self._f("set_invdt")(self.h, v)

    def make_boundaries(self, which: int):
        self._f("make_boundaries")(self.h, which)

    def godunov(self, nstep: int, dt: float):
        self._f("godunov")(self.h, nstep, dt)

    def run(self, nstepmax: int, tend: float = 10.0):
        dts = np.zeros(max(nstepmax, 1), dtype=np.float64)
        n = self._f("run")(self.h, nstepmax, tend, _dp(dts))
        return n, dts[:n]

    def get(self, which: int) -> np.ndarray:
        """Array as Fortran-ordered (isize, jsize, 4) float64."""
        out = np.empty(self.isize * self.jsize * 4, dtype=np.float64)
        self._f("get")(self.h, which, _dp(out))
        return out.reshape((self.isize, self.jsize, 4), order="F")

    def set(self, which: int, arr: np.ndarray):
        a = np.asfortranarray(arr, dtype=np.float64).reshape(-1, order="F").copy()
        self._f("set")(self.h, which, _dp(a))

    def fd(self, u0: np.ndarray, direction: np.ndarray, dts, eps: float, weight: np.ndarray | None = None,
           freeze: bool = False):
        """4-point central difference, in this oracle's precision, of the
        frozen-dt rollout U0 -> U_n (n = len(dts)) along `direction`: the
        (isize, jsize, 4) derivative of U_n, or with `weight` the derivative
        of sum(weight * U_n).  freeze: every face keeps the two-shock
        solver's decisions of the unperturbed run (Newton count, upwind side,
        shock or rarefaction, sampling).  u2 starts at zero; the run's
        arrays are overwritten."""
        a = np.asfortranarray(u0, dtype=np.float64).reshape(-1, order="F").copy()
        d = np.asfortranarray(direction, dtype=np.float64).reshape(-1, order="F").copy()
        t = np.ascontiguousarray(dts, dtype=np.float64)
        w = None if weight is None else np.asfortranarray(weight, dtype=np.float64).reshape(-1, order="F").copy()
        out = np.zeros(1 if w is not None else a.size, dtype=np.float64)
        rc = self._f("fd")(self.h, _dp(a), _dp(d), t.size, _dp(t), eps, None if w is None else _dp(w),
                           1 if freeze else 0, _dp(out))
        if rc != 0:
            raise MemoryError("oracle_fd scratch allocation failed")
        return float(out[0]) if w is not None else out.reshape((self.isize, self.jsize, 4), order="F")

    def kat(self, func: int, x: np.ndarray) -> np.ndarray:
        nin, nout = KAT_SHAPES[func]
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, nin)
        y = np.zeros((x.shape[0], nout), dtype=np.float64)
        self._f("kat")(self.h, func, x.shape[0], _dp(x), _dp(y))
        return y


def read_ref_dump(path) -> dict:
    """Parse oracle/ref_driver.f90 records -> {'dt': {step: v}, 'snap': {step: arr},
    'q','fx','fy','u_end','u2_end': arr}; arrays are flat float64 (Fortran order)."""
    out = {"dt": {}, "snap": {}}
    names = {3: "q", 4: "fx", 5: "fy", 6: "u_end", 7: "u2_end"}
    data = pathlib.Path(path).read_bytes()
    off = 0
    while off < len(data):
        tag, step, cnt = struct.unpack_from("<iiq", data, off)
        off += 16
        arr = np.frombuffer(data, dtype="<f8", count=cnt, offset=off).copy()
        off += 8 * cnt
        if tag == 1:
            out["dt"][step] = float(arr[0])
        elif tag == 2:
            out["snap"][step] = arr
        else:
            out[names[tag]] = arr
    return out


def write_nml(path, nx, ny=None, nstepmax=200, noutput=1000, tend=10.0, bc=(1, 1, 1, 1),
              gamma0=1.4, cfl=0.8, smallr=1e-7, smallc=1e-7, niter_riemann=10, iorder=2,
              slope_type=1, extra: str = ""):
    """Namelist file in the reference layout (serial_ref/test.nml) + extra groups."""
    ny = nx if ny is None else ny
    txt = (f"&RUN\ntEnd={tend}\nnStepmax={nstepmax}\nnOutput={noutput}\n/\n\n"
           f"&MESH\nnx={nx}\nny={ny}\nboundary_type_xmin={bc[0]}\nboundary_type_xmax={bc[1]}\n"
           f"boundary_type_ymin={bc[2]}\nboundary_type_ymax={bc[3]}\n/\n\n"
           f"&HYDRO\ngamma0={gamma0!r}\ncfl={cfl!r}\nsmallr={smallr!r}\nsmallc={smallc!r}\n"
           f"niter_riemann={niter_riemann}\niorder={iorder}\nslope_type={slope_type}\n/\n\n")
    pathlib.Path(path).write_text(txt + extra)
    return path


def rel_l1(a: np.ndarray, b: np.ndarray) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b).sum()
    return float(np.abs(a - b).sum() / den) if den > 0 else float(np.abs(a - b).sum())
