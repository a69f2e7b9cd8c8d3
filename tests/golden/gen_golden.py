"""Writes tests/golden/golden_small.npz and golden_small.json.

The reference repository holds no golden vectors or fixtures (its only check
is test_tridiagonal.cpp:107, RMS <= 1e-3 against its own serial Thomas), so
the fixtures here are produced independently of the oracle's Thomas code:

* inputs: the seeded generator, restated in numpy (oracle/oracle.py
  ``generate_np``), rounded to float32 or float64;
* expected solution: LAPACK banded LU with partial pivoting
  (``scipy.linalg.solve_banded``, i.e. ?gbsv) on the float64-promoted
  inputs -- a different algorithm from Thomas, from a different library;
* generator known answers: the first values of each stream for seed 0.

Run from the repository root:  python tests/golden/gen_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import scipy.linalg

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import generate_np, u01_np  # noqa: E402


def banded_solve(a, b, c, d, n):
    A, B, C, D = (np.asarray(v, dtype=np.float64).reshape(-1, n) for v in (a, b, c, d))
    out = np.empty_like(D)
    for s in range(A.shape[0]):
        ab = np.zeros((3, n))
        ab[0, 1:] = C[s, :-1]   # super-diagonal (c[n-1] ignored)
        ab[1, :] = B[s]
        ab[2, :-1] = A[s, 1:]   # sub-diagonal (a[0] ignored)
        out[s] = scipy.linalg.solve_banded((1, 1), ab, D[s])
    return out.reshape(-1)


def main():
    arrays = {}
    cases = []
    plan = []
    for n in (2, 4, 8, 16, 17, 64, 100, 512, 1000, 1024):
        N = 2 if n <= 100 else 1
        plan.append(("dominant", "float32", n, N))
        plan.append(("dominant", "float64", n, N))
    for n in (2, 17, 100, 1024):
        N = 2 if n <= 100 else 1
        plan.append(("weak", "float32", n, N))
        plan.append(("refgen", "float32", n, N))
    for i, (kind, dt, n, N) in enumerate(plan):
        seed = 1000 + i
        a, b, c, d = generate_np(kind, seed, n, N, np.dtype(dt))
        x = banded_solve(a, b, c, d, n)
        key = f"c{i}"
        arrays[key + "_a"], arrays[key + "_b"] = a, b
        arrays[key + "_c"], arrays[key + "_d"] = c, d
        arrays[key + "_x"] = x
        cases.append({"key": key, "kind": kind, "dtype": dt, "n": n, "N": N, "seed": seed})
    # generator known answers: seed 0, streams 0..3 (a,b,c,d), indices 0..7
    g = np.arange(8, dtype=np.uint64)
    arrays["kat_u01"] = np.stack([u01_np(0, s, g) for s in range(4)])
    ka, kb, kc, kd = generate_np("dominant", 0, 8, 1, np.float32)
    arrays["kat_dominant_f32"] = np.stack([ka, kb, kc, kd])
    np.savez_compressed(os.path.join(HERE, "golden_small.npz"), **arrays)
    with open(os.path.join(HERE, "golden_small.json"), "w") as f:
        json.dump({"generator": "tests/golden/gen_golden.py",
                   "expected": "scipy.linalg.solve_banded on float64-promoted inputs",
                   "cases": cases}, f, indent=1)
    print(f"{len(cases)} cases written")


if __name__ == "__main__":
    main()
