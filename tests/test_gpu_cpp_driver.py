"""Runs the C++ driver (tests/cpp/test_tridiagonal.cpp), which calls every
method through the reference's names (include/trid_compat.h) and checks each
system against the CPU serial Thomas with the reference's gate: RMS <= 1e-3
per system, NaN = failure (test_tridiagonal.cpp:107). Its exit status is the
failure count."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_tridiagonal")


@pytest.mark.parametrize("args", [("2", "8192", "f32"), ("16", "1000", "f64"),
                                  ("64", "512", "f32")])
def test_cpp_driver(args):
    assert os.path.exists(EXE), "built by __graft_entry__.build()"
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
