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


TABLE_EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_ragged_table")


@pytest.mark.parametrize("which", ["refine", "fp64"])
def test_cpp_ragged_plan_from_table_file(which, tmp_path):
    """A C++ host reads the shipped autotune table through
    trid_bucket_table_load and plans a ragged batch: the plan equals the one
    the Python binding builds from the same file (bucket configurations,
    systems per bucket, item counts, grid); in the FP64 mode the C++ host
    also solves it (relative residual <= 1e-12)."""
    import json

    import numpy as np
    import tridiagonalsolvers_amd.ragged as R
    assert os.path.exists(TABLE_EXE), "built by __graft_entry__.build()"
    path = R.DEFAULT_TABLE if which == "refine" else R.DEFAULT_TABLE_FP64
    hi = 16384 if which == "refine" else 8192
    sizes = np.concatenate([[2, 3, 512, 513, 8192, hi], R.log_uniform_sizes(3000, 2, hi, seed=3)])
    sp = tmp_path / "sizes.txt"
    sp.write_text(" ".join(str(int(v)) for v in sizes))
    args = [TABLE_EXE, path, str(sp)] + (["--fp64"] if which == "fp64" else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    got = json.loads(r.stdout.strip().splitlines()[-1])
    plan = R.RaggedPlan(R.offsets_of(sizes), table=R.load_table(path))
    ref = plan.info()
    plan.close()
    assert got["table_entries"] == len(R.load_table(path))
    for k in ("num_systems", "total_unknowns", "num_buckets", "workgroup_items",
              "wave_items", "grid"):
        assert got[k] == ref[k], k
    assert got["buckets"] == ref["buckets"]
    if which == "fp64":
        assert got["max_rel_residual"] <= 1e-12
