"""Adjoint building blocks on the CPU.

tests/adjoint_blocks_check.cpp compiles the device numerics and their
hand-written adjoints (euler2d_cudafortran_amd/csrc/e2d_adjoint.hpp) for the
host and checks, on random states, the dot-product identity
<J v, w> = <v, J^T w> against the dual-number tangent of the same forward
code, and the tangent against central finite differences.
"""
import re
import shutil
import subprocess

import pytest

from conftest import ROOT


@pytest.fixture(scope="module")
def report(tmp_path_factory):
    cxx = shutil.which("g++")
    assert cxx, "g++ not found"
    exe = tmp_path_factory.mktemp("adj") / "adjoint_blocks_check"
    subprocess.run([cxx, "-O2", "-std=c++17", "-ffp-contract=off", "-o", str(exe),
                    str(ROOT / "tests" / "adjoint_blocks_check.cpp")], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    rows = {}
    for line in out.splitlines():
        m = re.match(r"(\S+) n=(\d+) dot=(\S+) fd=(\S+)", line)
        rows[m.group(1)] = (int(m.group(2)), float(m.group(3)), float(m.group(4)))
    return rows


@pytest.mark.parametrize("name", ["riemann_two_shock", "riemann_y", "cons_to_prim", "limited_slope_st1",
                                  "limited_slope_st2", "trace"])
def test_dot_product_identity(report, name):
    n, dot, _ = report[name]
    assert n >= 4000
    assert dot <= 1e-10, f"{name}: <Jv,w> vs <v,J^T w> relative error {dot:.3e}"


@pytest.mark.parametrize("name", ["riemann_two_shock", "riemann_y"])
def test_tangent_against_finite_differences(report, name):
    _, _, fd = report[name]
    # worst case over 4000 random states, eps = 1e-6: states that sit within
    # eps of a branch switch (Newton count, sampling) dominate the maximum
    assert fd <= 1e-4, f"{name}: tangent vs finite difference {fd:.3e}"


def test_float_instantiation_against_double(report):
    # the FP32/mixed adjoint runs the same reverse code in float; on float inputs
    # it must agree with the double instantiation to FP32 rounding (99th
    # percentile; the tail is float forward passes taking another branch)
    n, p99, _ = report["riemann_two_shock_f32"]
    assert n >= 4000
    assert p99 <= 1e-4, f"float vs double Riemann adjoint, 99th percentile {p99:.3e}"
