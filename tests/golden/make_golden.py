"""Regenerate the golden fixtures under tests/golden/ from the compiled reference.

Test infrastructure.  Runs only where the reference sources are present: it
builds oracle/_ref (make -C oracle ref; the reference serial solver compiled
from serial_ref/ with flang, FP64 and FP32 kinds) and runs

  * oracle/_ref/<kind>/ref_kat   on seeded input batches -> kat_<name>_<kind>.npz
  * oracle/_ref/<kind>/ref_driver on small runs          -> run_<name>.npz
  * oracle/_ref/<kind>/euler2d   (the reference program, unmodified main.f90)
    on a 16x16 run, keeping its VTK files                -> vtk_<kind>_<step>.vti

Every fixture holds the inputs and the reference's outputs (data only), and
MANIFEST.sha256 lists them.  Usage:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import os
import pathlib
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT / "tests"))
from oracle_lib import read_ref_dump, write_nml  # noqa: E402

REF = ROOT / "oracle" / "_ref"
GAMMA, SMALLR, SMALLC = 1.4, 1e-7, 1e-7

# ---------------------------------------------------------------------------
# pointwise KAT inputs (seeded)


def _logu(rng, lo, hi, n):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n))


def _states(rng, n, rho=(1e-3, 10.0), p=(1e-3, 10.0), vel=5.0):
    """Primitive states (rho, p, u, v) with a share of special cases."""
    q = np.empty((n, 4))
    q[:, 0] = _logu(rng, *rho, n)
    q[:, 1] = _logu(rng, *p, n)
    q[:, 2] = rng.uniform(-vel, vel, n)
    q[:, 3] = rng.uniform(-vel, vel, n)
    k = n // 10
    q[:k, 0] = rng.uniform(0, 2e-7, k)           # density floor smallr
    q[k:2 * k, 1] = rng.uniform(-1e-3, 1e-12, k)  # pressure floor
    q[2 * k:3 * k, 2:] = 0.0                      # at rest
    return q


def kat_inputs(func: str, rng, n: int) -> tuple[np.ndarray, int, int]:
    """-> (x[n, nin], func_id, nout)"""
    if func == "prim":
        x = np.empty((n, 4))
        x[:, 0] = _logu(rng, 1e-9, 10.0, n)
        x[: n // 8, 0] = rng.uniform(-1e-3, 1e-7, n // 8)
        x[:, 2] = rng.uniform(-5, 5, n) * x[:, 0]
        x[:, 3] = rng.uniform(-5, 5, n) * x[:, 0]
        kin = 0.5 * (x[:, 2] ** 2 + x[:, 3] ** 2) / np.maximum(x[:, 0], SMALLR)
        x[:, 1] = kin + _logu(rng, 1e-6, 10.0, n)
        x[n // 2: n // 2 + n // 8, 1] = kin[n // 2: n // 2 + n // 8] * rng.uniform(0.5, 1.0, n // 8)  # e<0
        return x, 1, 5
    if func == "eos":
        x = np.empty((n, 2))
        x[:, 0] = _logu(rng, 1e-9, 10.0, n)
        x[:, 1] = rng.uniform(-1.0, 10.0, n)
        return x, 7, 2
    if func == "slope":
        base = _states(rng, n)
        x = np.empty((n, 20))
        x[:, 0:4] = base
        for k in range(4):
            x[:, 4 * (k + 1): 4 * (k + 2)] = base * (1 + rng.normal(0, 0.3, (n, 4))) + rng.normal(0, 0.2, (n, 4))
        m = n // 8
        x[:m, 4:8] = x[:m, 0:4]                     # drgt = 0 along x
        x[m:2 * m, 8:12] = x[m:2 * m, 0:4]          # dlft = 0
        x[2 * m:3 * m, 8:12] = x[2 * m:3 * m, 4:8]  # dcen = 0 along x
        x[3 * m:4 * m, 12:16] = x[3 * m:4 * m, 16:20]  # dcen = 0 along y
        mono = slice(4 * m, 5 * m)                  # monotone: q-x < q < q+x
        x[mono, 8:12] = x[mono, 0:4] - np.abs(rng.normal(0, 0.2, (m, 4)))
        x[mono, 4:8] = x[mono, 0:4] + np.abs(rng.normal(0, 0.2, (m, 4)))
        return x, 2, 8
    if func == "trace":
        x = np.empty((n, 15))
        x[:, 0:4] = _states(rng, n, rho=(1e-2, 10.0), p=(1e-2, 10.0))
        x[:, 4:12] = rng.normal(0, 0.3, (n, 8))
        x[:, 12] = rng.uniform(0, 0.5, n)
        x[:, 13] = rng.uniform(0, 0.5, n)
        x[:, 14] = rng.integers(0, 4, n)
        return x, 3, 4
    if func in ("riemann", "hllc", "riemann_floor"):
        ql = _states(rng, n)
        qr = _states(rng, n)
        m = n // 10
        qr[:m] = ql[:m]                                   # identical states
        qr[m:2 * m, 0:2] = ql[m:2 * m, 0:2]
        qr[m:2 * m, 2] = -ql[m:2 * m, 2]                  # symmetric collision/expansion
        ql[2 * m:3 * m, 2] = -rng.uniform(5, 50, m)       # strong rarefaction
        qr[2 * m:3 * m, 2] = rng.uniform(5, 50, m)
        ql[3 * m:4 * m, 2] = rng.uniform(5, 50, m)        # strong collision
        qr[3 * m:4 * m, 2] = -rng.uniform(5, 50, m)
        ql[4 * m:5 * m, 1] = _logu(rng, 1e2, 1e5, m)      # strong pressure jump
        x = np.concatenate([ql, qr], axis=1)
        return x, (5 if func == "hllc" else 4), 8
    if func == "cmpflx":
        x = _states(rng, n)
        return x, 6, 4
    raise ValueError(func)


KAT_SETS = [
    # name, func, n, slope_type, niter, gamma0, smallr, smallc
    ("prim", "prim", 2000, 1, 10, GAMMA, SMALLR, SMALLC),
    ("eos", "eos", 1000, 1, 10, GAMMA, SMALLR, SMALLC),
    ("slope_st0", "slope", 500, 0, 10, GAMMA, SMALLR, SMALLC),
    ("slope_st1", "slope", 2000, 1, 10, GAMMA, SMALLR, SMALLC),
    ("slope_st2", "slope", 2000, 2, 10, GAMMA, SMALLR, SMALLC),
    ("trace", "trace", 2000, 1, 10, GAMMA, SMALLR, SMALLC),
    ("riemann", "riemann", 4000, 1, 10, GAMMA, SMALLR, SMALLC),
    ("riemann_niter1", "riemann", 1000, 1, 1, GAMMA, SMALLR, SMALLC),
    ("riemann_floor", "riemann_floor", 1000, 1, 10, 1.666, 1e-2, 1e-2),
    ("hllc", "hllc", 4000, 1, 10, GAMMA, SMALLR, SMALLC),
    ("cmpflx", "cmpflx", 1000, 1, 10, GAMMA, SMALLR, SMALLC),
]


def run_kat(kind: str, func_id: int, x: np.ndarray, nout: int, slope_type: int, niter: int,
            gamma0: float, smallr: float, smallc: float, tmp: pathlib.Path) -> np.ndarray:
    n, nin = x.shape
    hdr = np.array([func_id, n, slope_type, niter, nin, nout, 0, 0], dtype="<i4")
    par = np.array([gamma0, smallr, smallc, 0, 0, 0, 0, 0], dtype="<f8")
    fin, fout = tmp / "kat_in.bin", tmp / "kat_out.bin"
    with open(fin, "wb") as f:
        f.write(hdr.tobytes())
        f.write(par.tobytes())
        f.write(np.ascontiguousarray(x, dtype="<f8").tobytes())
    subprocess.run([str(REF / kind / "ref_kat"), str(fin), str(fout)], check=True)
    return np.frombuffer(fout.read_bytes(), dtype="<f8").reshape(n, nout).copy()


# ---------------------------------------------------------------------------
# whole-run fixtures

RUN_SETS = [
    # name, kind, nx, ny, ic, bc, slope_type, gamma0, steps, dump_steps, stages
    ("diag128_fp64", "fp64", 128, 128, "diag", (1, 1, 1, 1), 1, 1.4, 200, (0, 1, 9), True),
    ("implode128_fp64", "fp64", 128, 128, "implode", (1, 1, 1, 1), 1, 1.4, 200, (), False),
    ("diag128_fp32", "fp32", 128, 128, "diag", (1, 1, 1, 1), 1, 1.4, 200, (), False),
    ("implode64_fp32_mc", "fp32", 64, 64, "implode", (1, 1, 1, 1), 2, 1.666, 100, (), False),
    ("diag48x40_fp64_absorb_mc", "fp64", 48, 40, "diag", (2, 2, 2, 2), 2, 1.4, 80, (), True),
    ("implode40x48_fp64_periodic", "fp64", 40, 48, "implode", (3, 3, 3, 3), 1, 1.4, 80, (), False),
    ("diag32_fp64_mixedbc_st0", "fp64", 32, 32, "diag", (1, 2, 3, 1), 0, 1.4, 60, (), False),
    ("implode50x30_fp64_mixedbc", "fp64", 50, 30, "implode", (1, 1, 2, 3), 2, 1.666, 60, (), False),
]


def run_fixture(name, kind, nx, ny, ic, bc, st, g, steps, dumps, stages, tmp):
    nml = tmp / f"{name}.nml"
    dump = tmp / f"{name}.bin"
    ds = ",".join(str(s) for s in dumps) if dumps else "-1"
    write_nml(nml, nx, ny, nstepmax=steps, bc=bc, gamma0=g, slope_type=st,
              extra=(f"&ORACLE\ninitial_condition='{ic}'\ndump_file='{dump}'\n"
                     f"dump_steps={ds}\ndump_stages={'.true.' if stages else '.false.'}\n/\n"))
    subprocess.run([str(REF / kind / "ref_driver"), str(nml)], check=True)
    d = read_ref_dump(dump)
    sdt = np.float32 if kind == "fp32" else np.float64
    shape = (nx + 4, ny + 4, 4)
    arrs = {
        "params": np.array([nx, ny, *bc, st, steps, 1 if ic == "diag" else 0], dtype=np.int64),
        "gamma0": np.float64(g),
        "dt": np.array([d["dt"][k] for k in range(steps)], dtype=np.float64),
        "u_end": d["u_end"].astype(sdt).reshape(shape, order="F"),
        "u2_end": d["u2_end"].astype(sdt).reshape(shape, order="F"),
    }
    for s, a in d["snap"].items():
        arrs[f"snap_{s}"] = a.astype(sdt).reshape(shape, order="F")
    for k in ("q", "fx", "fy"):
        if k in d:
            arrs[f"stage0_{k}"] = d[k].astype(sdt).reshape(shape, order="F")
    np.savez_compressed(HERE / f"run_{name}.npz", **arrs)


def vtk_fixture(kind, tmp):
    """The reference program itself (serial_ref/main.f90), 16x16, 3 steps, nOutput=2."""
    wd = tmp / f"vtk_{kind}"
    wd.mkdir()
    write_nml(wd / "test.nml", 16, 16, nstepmax=3, noutput=2, gamma0=1.4, slope_type=1)
    subprocess.run([str(REF / kind / "euler2d"), "test.nml"], check=True, cwd=wd,
                   stdout=subprocess.DEVNULL)
    for f in sorted(wd.glob("euler2d_*.vti")):
        shutil.copy(f, HERE / f"vtk_{kind}_{f.stem.split('_')[1]}.vti")


def main():
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "all", "ref"], check=True)
    rng_seed = 20251202
    with tempfile.TemporaryDirectory() as td:
        tmp = pathlib.Path(td)
        for kind in ("fp64", "fp32"):
            for i, (name, func, n, st, niter, g, sr, sc) in enumerate(KAT_SETS):
                rng = np.random.default_rng(rng_seed + 97 * i)
                x, fid, nout = kat_inputs(func, rng, n)
                if kind == "fp32":
                    x = x.astype(np.float32).astype(np.float64)
                y = run_kat(kind, fid, x, nout, st, niter, g, sr, sc, tmp)
                np.savez_compressed(HERE / f"kat_{name}_{kind}.npz", x=x, y=y, func=fid,
                                    slope_type=st, niter=niter, gamma0=g, smallr=sr, smallc=sc)
        for spec in RUN_SETS:
            run_fixture(*spec, tmp)
        for kind in ("fp64", "fp32"):
            vtk_fixture(kind, tmp)
    lines = []
    for f in sorted(HERE.iterdir()):
        if f.suffix in (".npz", ".vti"):
            lines.append(f"{hashlib.sha256(f.read_bytes()).hexdigest()}  {f.name}")
    (HERE / "MANIFEST.sha256").write_text("\n".join(lines) + "\n")
    print("\n".join(lines))


if __name__ == "__main__":
    main()
