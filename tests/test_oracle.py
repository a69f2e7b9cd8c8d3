"""Pins the CPU oracle (oracle/) before anything is checked against it.

The reference has no golden vectors (its one check is test_tridiagonal.cpp:107),
so the oracle is pinned by: the committed fixtures (LAPACK banded LU, an
independent algorithm), closed forms, manufactured solutions, numpy dense
solves, an independent numpy restatement, and planted mistakes that these
checks must catch.
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def load_golden():
    meta = json.load(open(os.path.join(GOLDEN, "golden_small.json")))
    data = np.load(os.path.join(GOLDEN, "golden_small.npz"))
    return meta, data


def relerr(x, ref, n):
    x = np.asarray(x, dtype=np.float64).reshape(-1, n)
    ref = np.asarray(ref, dtype=np.float64).reshape(-1, n)
    return float(np.max(np.abs(x - ref).max(axis=1) / np.abs(ref).max(axis=1)))


# --------------------------------------------------------------------------
# generator
# --------------------------------------------------------------------------

@pytest.mark.parametrize("kind", ["dominant", "weak", "refgen", "paired", "general"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_generator_c_equals_numpy(oracle, kind, dtype):
    for n, N in ((2, 3), (17, 5), (1024, 4)):
        c_arrs = oracle.generate(kind, 7, n, N, dtype)
        np_arrs = oracle.generate_np(kind, 7, n, N, dtype)
        for u, v in zip(c_arrs, np_arrs):
            assert u.dtype == np.dtype(dtype)
            assert np.array_equal(u.view(np.uint8), v.view(np.uint8))


def test_generator_ragged_c_equals_numpy(oracle):
    offsets = np.array([0, 5, 11, 111, 113], dtype=np.int64)
    for u, v in zip(oracle.generate("dominant", 3, 0, 0, np.float32, offsets=offsets),
                    oracle.generate_np("dominant", 3, 0, 0, np.float32, offsets=offsets)):
        assert np.array_equal(u, v)


def test_generator_known_answers(oracle):
    _, data = load_golden()
    kat = data["kat_u01"]
    for s in range(4):
        for g in range(8):
            assert oracle.lib().oracle_u01(0, s, g) == kat[s, g]
    a, b, c, d = oracle.generate("dominant", 0, 8, 1, np.float32)
    assert np.array_equal(np.stack([a, b, c, d]), data["kat_dominant_f32"])


@pytest.mark.parametrize("kind", ["dominant", "weak", "refgen"])
def test_generator_properties(oracle, kind):
    n, N = 64, 32
    a, b, c, d = (v.reshape(N, n).astype(np.float64)
                  for v in oracle.generate(kind, 11, n, N, np.float32))
    assert np.all(a[:, 0] == 0) and np.all(c[:, -1] == 0)
    margin = b - np.abs(a) - np.abs(c)
    if kind == "dominant":
        assert margin.min() >= 1.0 - 1e-6 and margin.max() <= 2.0 + 1e-6
        assert a.min() >= -1 and a.max() <= 1 and d.min() >= -1 and d.max() <= 1
    elif kind == "weak":
        assert margin.min() >= 0.01 - 1e-6 and margin.max() <= 0.1 + 1e-6
    else:
        assert 8 <= b.min() and b.max() <= 9
        assert 3 <= a[:, 1:].min() and a.max() <= 4
        assert 2 <= c[:, :-1].min() and c.max() <= 3
        assert 5 <= d.min() and d.max() <= 6
    # different seeds give different streams, same seed the same stream
    a2 = oracle.generate(kind, 12, n, N, np.float32)[0]
    a3 = oracle.generate(kind, 11, n, N, np.float32)[0]
    assert not np.array_equal(a2, a.reshape(-1).astype(np.float32))
    assert np.array_equal(a3, a.reshape(-1).astype(np.float32))


# --------------------------------------------------------------------------
# Thomas
# --------------------------------------------------------------------------

def test_thomas_against_golden(oracle):
    meta, data = load_golden()
    for case in meta["cases"]:
        k, n = case["key"], case["n"]
        a, b, c, d, x = (data[f"{k}_{s}"] for s in "abcdx")
        xo = oracle.thomas(a, b, c, d, n=n)
        tol = 1e-12 if case["kind"] != "weak" else 1e-10
        assert relerr(xo, x, n) <= tol, case
        # the inputs in the fixture are exactly what the generator makes
        g = oracle.generate(case["kind"], case["seed"], n, case["N"], np.dtype(case["dtype"]))
        for u, v in zip(g, (a, b, c, d)):
            assert np.array_equal(u, v)


def test_thomas_closed_form_diagonal(oracle):
    n, N = 50, 3
    rng = np.random.default_rng(0)
    b = rng.uniform(1, 2, n * N)
    d = rng.uniform(-1, 1, n * N)
    z = np.zeros(n * N)
    x = oracle.thomas(z, b, z, d, n=n)
    assert np.array_equal(x, d / b)


@pytest.mark.parametrize("n", [2, 3, 10, 100])
def test_thomas_closed_form_poisson(oracle, n):
    # tridiag(-1, 2, -1) x = 1  =>  x_i = i (n + 1 - i) / 2, i = 1..n
    a = -np.ones(n)
    c = -np.ones(n)
    b = 2 * np.ones(n)
    d = np.ones(n)
    a[0] = 0
    c[-1] = 0
    i = np.arange(1, n + 1, dtype=np.float64)
    exact = i * (n + 1 - i) / 2
    x = oracle.thomas(a, b, c, d, n=n)
    assert relerr(x, exact, n) <= 1e-13


def test_thomas_ignores_first_a_and_last_c(oracle):
    n, N = 16, 4
    a, b, c, d = oracle.generate("dominant", 5, n, N, np.float64)
    x1 = oracle.thomas(a, b, c, d, n=n)
    a2, c2 = a.copy(), c.copy()
    a2[::n] = 123.0
    c2[n - 1::n] = -77.0
    x2 = oracle.thomas(a2, b, c2, d, n=n)
    assert np.array_equal(x1, x2)


def test_thomas_manufactured_solution(oracle):
    n, N = 257, 8
    a, b, c, _ = oracle.generate("weak", 9, n, N, np.float64)
    rng = np.random.default_rng(1)
    xs = rng.uniform(-1, 1, n * N)
    A, B, C, X = (v.reshape(N, n) for v in (a, b, c, xs))
    D = B * X
    D[:, 1:] += A[:, 1:] * X[:, :-1]
    D[:, :-1] += C[:, :-1] * X[:, 1:]
    x = oracle.thomas(a, b, c, D.reshape(-1), n=n)
    assert relerr(x, xs, n) <= 1e-12


@pytest.mark.parametrize("n", [2, 5, 32, 64])
def test_thomas_dense_solve(oracle, n):
    a, b, c, d = oracle.generate("dominant", 21, n, 3, np.float64)
    x = oracle.thomas(a, b, c, d, n=n).reshape(3, n)
    for s in range(3):
        T = np.diag(b[s * n:(s + 1) * n])
        T += np.diag(a[s * n + 1:(s + 1) * n], -1)
        T += np.diag(c[s * n:(s + 1) * n - 1], 1)
        ref = np.linalg.solve(T, d[s * n:(s + 1) * n])
        assert np.max(np.abs(x[s] - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_thomas_numpy_restatement_agrees(oracle):
    n, N = 512, 16
    a, b, c, d = oracle.generate("dominant", 2, n, N, np.float32)
    x_c = oracle.thomas(a, b, c, d, n=n)
    x_np = oracle.thomas_np(a, b, c, d, n)
    assert relerr(x_c, x_np, n) <= 1e-15


def test_thomas_ragged_matches_uniform(oracle):
    sizes = [2, 17, 64, 5, 300]
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    a, b, c, d = oracle.generate("dominant", 4, 0, 0, np.float32, offsets=offsets)
    x = oracle.thomas(a, b, c, d, offsets=offsets)
    for s, n in enumerate(sizes):
        o = offsets[s]
        xs = oracle.thomas(a[o:o + n], b[o:o + n], c[o:o + n], d[o:o + n], n=n)
        assert np.array_equal(x[o:o + n], xs)


def test_thomas_f32_close_to_f64(oracle):
    n, N = 1024, 32
    a, b, c, d = oracle.generate("dominant", 8, n, N, np.float32)
    x64 = oracle.thomas(a, b, c, d, n=n)
    x32 = oracle.thomas_f32(a, b, c, d, n)
    assert x32.dtype == np.float32
    assert relerr(x32, x64, n) <= 1e-5


def test_thomas_all_systems_solved_any_count(oracle):
    # the reference's 4-thread driver drops numSystems mod 4 systems
    # (tridiagonal_gold.h:58-67); the oracle must not
    n = 8
    for N in (1, 3, 5, 7):
        a, b, c, d = oracle.generate("dominant", 1, n, N, np.float64)
        x = oracle.thomas(a, b, c, d, n=n, threads=4)
        ref = oracle.thomas_np(a, b, c, d, n)
        assert np.allclose(x, ref, rtol=1e-14, atol=0)


# --------------------------------------------------------------------------
# planted mistakes: each plausible slip in a Thomas restatement must fail the
# golden check (so the golden check is able to catch them in the oracle)
# --------------------------------------------------------------------------

def _thomas_variant(a, b, c, d, n, mistake):
    A, B, C, D = (np.asarray(v, dtype=np.float64).reshape(-1, n).copy() for v in (a, b, c, d))
    if mistake == "swap_ac":
        A, C = C.copy(), A.copy()
        A[:, 0] = 0
        C[:, -1] = 0
    C[:, -1] = 0
    N = A.shape[0]
    cp = np.zeros((N, n))
    dp = np.zeros((N, n))
    cp[:, 0] = C[:, 0] / B[:, 0]
    dp[:, 0] = D[:, 0] / B[:, 0]
    for i in range(1, n):
        ai = A[:, i - 1] if mistake == "shift_a" else A[:, i]
        m = B[:, i] - ai * cp[:, i - 1]
        if mistake == "sign":
            m = B[:, i] + ai * cp[:, i - 1]
        cp[:, i] = C[:, i] / m
        dp[:, i] = (D[:, i] - dp[:, i - 1] * ai) / m
    x = np.zeros((N, n))
    x[:, -1] = dp[:, -1]
    for i in range(n - 2, -1, -1):
        if mistake == "drop_back":
            x[:, i] = dp[:, i]
        else:
            x[:, i] = dp[:, i] - cp[:, i] * x[:, i + 1]
    return x.reshape(-1)


@pytest.mark.parametrize("mistake", ["none", "swap_ac", "shift_a", "sign", "drop_back"])
def test_planted_mistakes_are_caught(mistake):
    meta, data = load_golden()
    worst = 0.0
    for case in meta["cases"]:
        if case["n"] < 8:
            continue
        k, n = case["key"], case["n"]
        a, b, c, d, x = (data[f"{k}_{s}"] for s in "abcdx")
        worst = max(worst, relerr(_thomas_variant(a, b, c, d, n, mistake), x, n))
    if mistake == "none":
        assert worst <= 1e-10
    else:
        assert worst > 1e-3


# --------------------------------------------------------------------------
# metrics
# --------------------------------------------------------------------------

def test_compare_rms_and_nan(oracle):
    ref = np.array([1.0, 2.0, 3.0, 4.0, 1.0, 1.0, 1.0, 1.0])
    x = ref.copy()
    x[0] += 0.0019          # system 0: rms = 0.00095 -> passes at eps 1e-3
    x[5] = np.nan           # system 1: NaN -> fails
    rel, rms, fails = oracle.compare(x, ref, n=4, rms_eps=1e-3)
    assert abs(rms[0] - 0.00095) < 1e-12
    assert abs(rel[0] - 0.0019 / 4.0) < 1e-12
    assert np.isnan(rms[1]) and np.isnan(rel[1])
    assert fails == 1
    _, _, fails2 = oracle.compare(x, ref, n=4, rms_eps=0.0009)
    assert fails2 == 2


def test_residual(oracle):
    n, N = 128, 4
    a, b, c, d = oracle.generate("dominant", 3, n, N, np.float64)
    x = oracle.thomas(a, b, c, d, n=n)
    assert oracle.residual(a, b, c, d, x, n=n).max() <= 1e-14
    x2 = x.copy()
    x2[5] += 1e-6
    r = oracle.residual(a, b, c, d, x2, n=n)
    assert r[0] > 1e-7 and r[1:].max() <= 1e-14


@pytest.mark.parametrize("n", [7, 32])
def test_adi_oracle_mode_closed_form(oracle, n):
    # the numpy ADI step against the exact amplification of a sine mode
    i = np.arange(1, n + 1)
    for p, q, r in ((1, 1, 0.3), (3, 5, 2.0), (n, 2, 9.5)):
        u0 = np.outer(np.sin(p * np.pi * i / (n + 1)), np.sin(q * np.pi * i / (n + 1)))
        u1 = oracle.adi_step_np(u0, r)
        g = oracle.adi_mode_factor(n, p, q, r)
        assert np.max(np.abs(u1 - g * u0)) <= 1e-13 * max(1.0, abs(g))


def _pcr_split_solve(oracle, a, b, c, d, n, G, levels):
    """numpy restatement of `levels` PCR levels (pcr_kernel.cu:39-56, identity
    rows outside the system), then Thomas on each subsystem {r, r+G, ...}."""
    A, B, C, D = (v.astype(np.float64).copy() for v in (a, b, c, d))
    A[0] = 0.0
    C[-1] = 0.0
    delta = 1
    for _ in range(levels):
        idx = np.arange(n)
        lo, hi = idx - delta, idx + delta
        okl, okh = lo >= 0, hi < n
        li, hj = np.clip(lo, 0, n - 1), np.clip(hi, 0, n - 1)
        aL, bL = np.where(okl, A[li], 0.0), np.where(okl, B[li], 1.0)
        cL, dL = np.where(okl, C[li], 0.0), np.where(okl, D[li], 0.0)
        aR, bR = np.where(okh, A[hj], 0.0), np.where(okh, B[hj], 1.0)
        cR, dR = np.where(okh, C[hj], 0.0), np.where(okh, D[hj], 0.0)
        k1, k2 = A / bL, C / bR
        A, B, C, D = -aL * k1, B - cL * k1 - aR * k2, -cR * k2, D - dL * k1 - dR * k2
        delta *= 2
    x = np.empty(n)
    for r in range(min(G, n)):
        rows = np.arange(r, n, G)
        x[rows] = oracle.thomas(A[rows], B[rows], C[rows], D[rows], n=len(rows))
    return x


@pytest.mark.parametrize("n,G", [(64, 4), (100, 8), (1024, 64), (37, 16)])
def test_pcr_levels_split_into_independent_subsystems(oracle, n, G):
    # the math pcrGauss rests on: after log2(G) PCR levels equation i couples
    # only to i -+ G, so the subsystems {r, r+G, ...} solved on their own give
    # the full solution; one level short, they do not
    a, b, c, d = oracle.generate("weak", 9, n, 1, np.float64)
    xref = oracle.thomas(a, b, c, d, n=n)
    depth = G.bit_length() - 1
    x = _pcr_split_solve(oracle, a, b, c, d, n, G, depth)
    err = np.max(np.abs(x - xref)) / np.max(np.abs(xref))
    assert err <= 1e-10
    # (the coupling left after one level too few decays fast on dominant
    # systems, but stays far above rounding)
    x_short = _pcr_split_solve(oracle, a, b, c, d, n, G, depth - 1)
    err_short = np.max(np.abs(x_short - xref)) / np.max(np.abs(xref))
    assert err_short > max(100 * err, 1e-12), (err, err_short)
