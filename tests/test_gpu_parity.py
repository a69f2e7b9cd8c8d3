"""GPU parity of the register-resident CR / CR-PCR / PCR kernels against the
FP64 Thomas oracle (gold recurrence of tridiagonal_gold.h:26-44).

Tolerances (max over systems of ||x - x_thomas||_inf / ||x_thomas||_inf on
diagonally dominant inputs): FP32 <= 1e-4, FP64 <= 1e-11. The reference's
own gate (RMS <= 1e-3 per system, NaN = failure, test_tridiagonal.cpp:107)
is checked too.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = {np.float32: 1e-4, np.float64: 1e-11}
HERE = os.path.dirname(os.path.abspath(__file__))


def solve_host(T, method, a, b, c, d, n, N, param=None):
    x = np.full_like(a, np.nan)
    f = {"cr": T.cr, "pcr": T.pcr, "crnbc": T.crNBC, "crnbc32": T.crNBC32}.get(method)
    if f is not None:
        f(a, b, c, d, x, n, N)
    elif method == "crpcr":
        T.crpcr(a, b, c, d, x, n, N, param)
    elif method == "crpcrnbc":
        T.crpcrNBC(a, b, c, d, x, n, N, param)
    elif method == "pcrgauss":
        T.pcrGaussWrapper(a, b, c, d, x, n, N, param)
    return x


def check(O, x, a, b, c, d, n, dtype, tol=None):
    xref = O.thomas(a, b, c, d, n=n)
    rel, rms, fails = O.compare(x, xref, n=n, rms_eps=1e-3)
    worst = float(np.max(rel))
    tol = TOL[dtype] if tol is None else tol
    assert worst <= tol, f"max rel err {worst:.3e} > {tol:.1e}"
    assert fails == 0
    return worst


SIZES = [2, 4, 8, 16, 17, 31, 64, 100, 128, 256, 512, 1000, 1024]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("method", ["cr", "pcr", "crpcr"])
def test_register_methods_sizes(oracle, trid, method, dtype):
    for n in SIZES:
        N = 37
        a, b, c, d = oracle.generate("dominant", 100 + n, n, N, dtype)
        rest = None
        if method == "crpcr":
            n_pad = 1 << (n - 1).bit_length()
            if n_pad < 4:
                # no switch size in [2, n/2] exists (crpcr_app.cu:5 contract)
                with pytest.raises(trid.TridError):
                    solve_host(trid, method, a, b, c, d, n, N, 2)
                continue
            rest = min(64, n_pad // 2)
        x = solve_host(trid, method, a, b, c, d, n, N, rest)
        print(f"{method} {np.dtype(dtype).name} n={n} N={N}", flush=True)
        check(oracle, x, a, b, c, d, n, dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_crpcr_rest_sweep_agrees(oracle, trid, dtype):
    n, N = 1024, 64
    a, b, c, d = oracle.generate("dominant", 7, n, N, dtype)
    results = {}
    rest = 2
    while rest <= n // 2:
        x = solve_host(trid, "crpcr", a, b, c, d, n, N, rest)
        check(oracle, x, a, b, c, d, n, dtype)
        results[rest] = x
        rest *= 2
    base = results[2].astype(np.float64)
    for r, x in results.items():
        assert np.max(np.abs(x - base)) <= 10 * TOL[dtype] * np.max(np.abs(base)), r


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_lane_shapes_agree(oracle, trid, dtype):
    import torch
    n, N = 64, 50
    a, b, c, d = oracle.generate("dominant", 3, n, N, dtype)
    h = trid.Handle()
    dev = [torch.from_numpy(v).cuda() for v in (a, b, c, d)]
    for M in (1, 2, 4, 8, 16):
        L = n // M
        if L < 2 or L > 64:
            continue
        for rest in (2, 4, 8, 16, 32):
            x = torch.full_like(dev[0], float("nan"))
            h.solve("crpcr", *dev, x, n, N, trid.make_opts(rest=rest, lanes=L, per_lane=M))
            torch.cuda.synchronize()
            check(oracle, x.cpu().numpy(), a, b, c, d, n, dtype)
    h.close()


@pytest.mark.parametrize("kind", ["weak", "refgen"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_other_generators(oracle, trid, kind, dtype):
    n, N = 1024, 128
    a, b, c, d = oracle.generate(kind, 5, n, N, dtype)
    for method in ("cr", "pcr", "crpcr"):
        x = solve_host(trid, method, a, b, c, d, n, N, 64)
        # the weakly dominant variant (margin U[0.01,0.1]) is worse conditioned
        tol = TOL[dtype] * (10 if kind == "weak" else 1)
        worst = check(oracle, x, a, b, c, d, n, dtype, tol)
        print(f"{kind} {method} {np.dtype(dtype).name}: {worst:.2e}", flush=True)


def test_golden_fixtures(trid):
    meta = json.load(open(os.path.join(HERE, "golden", "golden_small.json")))
    data = np.load(os.path.join(HERE, "golden", "golden_small.npz"))
    for case in meta["cases"]:
        k, n, N = case["key"], case["n"], case["N"]
        a, b, c, d, xg = (data[f"{k}_{s}"] for s in "abcdx")
        dtype = a.dtype.type
        for method in ("cr", "pcr", "crpcr"):
            rest = min(64, (1 << (n - 1).bit_length()) // 2)
            if method == "crpcr" and rest < 2:
                continue
            x = solve_host(trid, method, a, b, c, d, n, N, rest)
            X, XG = x.reshape(N, n).astype(np.float64), xg.reshape(N, n)
            rel = np.max(np.abs(X - XG).max(axis=1) / np.abs(XG).max(axis=1))
            tol = TOL[dtype] * (10 if case["kind"] == "weak" else 1)
            assert rel <= tol, (case, method, rel)


def test_first_a_last_c_ignored(oracle, trid):
    n, N = 256, 9
    a, b, c, d = oracle.generate("dominant", 9, n, N, np.float32)
    x1 = solve_host(trid, "crpcr", a, b, c, d, n, N, 32)
    a2, c2 = a.copy(), c.copy()
    a2[::n] = 1e3
    c2[n - 1::n] = -1e3
    x2 = solve_host(trid, "crpcr", a2, b, c2, d, n, N, 32)
    assert np.array_equal(x1, x2)


def test_batch_position_invariance(oracle, trid):
    # each system's result is bit-identical alone and inside any batch
    n, N = 1024, 40
    for dtype in (np.float32, np.float64):
        a, b, c, d = oracle.generate("dominant", 13, n, N, dtype)
        x = solve_host(trid, "crpcr", a, b, c, d, n, N, 64)
        perm = np.random.default_rng(0).permutation(N)
        P = [v.reshape(N, n)[perm].reshape(-1).copy() for v in (a, b, c, d)]
        xp = solve_host(trid, "crpcr", *P, n, N, 64)
        assert np.array_equal(xp.reshape(N, n), x.reshape(N, n)[perm])
        for s in (0, 17, N - 1):
            sl = slice(s * n, (s + 1) * n)
            xs = solve_host(trid, "crpcr", a[sl].copy(), b[sl].copy(), c[sl].copy(),
                            d[sl].copy(), n, 1, 64)
            assert np.array_equal(xs, x[sl])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [64, 8192])
def test_zero_pivot_gives_nonfinite_not_crash(oracle, trid, dtype, n):
    N = 4
    a, b, c, d = oracle.generate("dominant", 2, n, N, dtype)
    b[n + 5] = 0.0
    a[n + 5] = 0.0
    c[n + 5] = 0.0   # system 1 is singular
    x = solve_host(trid, "crpcr", a, b, c, d, n, N, 32)
    X = x.reshape(N, n)
    assert not np.all(np.isfinite(X[1]))
    check(oracle, np.concatenate([X[0], X[2], X[3]]),
          *(np.concatenate([v.reshape(N, n)[i] for i in (0, 2, 3)]) for v in (a, b, c, d)),
          n, dtype)


def test_device_generator_bitwise_equals_oracle(oracle, trid):
    import torch
    for dtype, tdt in ((np.float32, torch.float32), (np.float64, torch.float64)):
        for kind, kname in ((0, "dominant"), (1, "weak"), (2, "refgen"), (3, "paired"),
                            (4, "general")):
            n, N = 1000, 33
            dev = [torch.empty(n * N, dtype=tdt, device="cuda") for _ in range(4)]
            trid.generate(kind, 42, *dev, n, N)
            torch.cuda.synchronize()
            ref = oracle.generate(kname, 42, n, N, dtype)
            for u, v in zip(dev, ref):
                assert np.array_equal(u.cpu().numpy(), v), (kname, dtype)


def test_device_api_matches_host_dropin(oracle, trid):
    import torch
    n, N = 1024, 256
    a, b, c, d = oracle.generate("dominant", 21, n, N, np.float32)
    x_host = solve_host(trid, "crpcr", a, b, c, d, n, N, 64)
    dev = [torch.from_numpy(v).cuda() for v in (a, b, c, d)]
    x = torch.empty_like(dev[0])
    s = torch.cuda.Stream()
    h = trid.Handle()
    with torch.cuda.stream(s):
        h.solve("crpcr", *dev, x, n, N, trid.make_opts(rest=64), stream=s)
    s.synchronize()
    assert np.array_equal(x.cpu().numpy(), x_host)
    h.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_65536x1024_full_size(oracle, trid, dtype):
    # 65,536 systems x 1,024 unknowns, whole batch checked
    import torch
    n, N = 1024, 65536
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    dev = [torch.empty(n * N, dtype=tdt, device="cuda") for _ in range(4)]
    trid.generate(0, 1, *dev, n, N)
    x = torch.empty_like(dev[0])
    h = trid.Handle()
    h.solve("crpcr", *dev, x, n, N, trid.make_opts(rest=64))
    torch.cuda.synchronize()
    a, b, c, d = oracle.generate("dominant", 1, n, N, dtype)
    worst = check(oracle, x.cpu().numpy(), a, b, c, d, n, dtype)
    print(f"65536x1024 {np.dtype(dtype).name}: max rel err {worst:.3e}", flush=True)
    h.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_io_modes_bitwise_identical(oracle, trid, dtype):
    # the dense-path I/O modes move the same values: results are bit-identical
    import torch
    for n, N in ((1024, 301), (512, 77), (64, 1000), (16, 129)):
        a, b, c, d = oracle.generate("dominant", 31, n, N, dtype)
        dev = [torch.from_numpy(v).cuda() for v in (a, b, c, d)]
        h = trid.Handle()
        outs = []
        for io in (1, 2, 3, 4):
            x = torch.full_like(dev[0], float("nan"))
            h.solve("crpcr", *dev, x, n, N, trid.make_opts(rest=min(64, n // 2), io_mode=io))
            torch.cuda.synchronize()
            outs.append(x.cpu().numpy())
        h.close()
        check(oracle, outs[0], a, b, c, d, n, dtype)
        for o in outs[1:]:
            assert np.array_equal(o, outs[0]), (n, N)


def _mw_sizes(dtype):
    return [2048, 3000, 4096, 8192] + ([16384] if dtype == np.float32 else [5000])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("method", ["cr", "crpcr"])
def test_multiwave_methods(oracle, trid, method, dtype):
    # systems longer than one wave: W = 2..16 waves per workgroup
    for n in _mw_sizes(dtype):
        N = 19
        a, b, c, d = oracle.generate("dominant", 200 + n, n, N, dtype)
        print(f"{method} {np.dtype(dtype).name} n={n}", flush=True)
        x = solve_host(trid, method, a, b, c, d, n, N, 64)
        check(oracle, x, a, b, c, d, n, dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_multiwave_rest_sweep(oracle, trid, dtype):
    import torch
    n, N = 8192, 24
    a, b, c, d = oracle.generate("dominant", 17, n, N, dtype)
    dev = [torch.from_numpy(v).cuda() for v in (a, b, c, d)]
    h = trid.Handle()
    lanes = trid.plan_query("crpcr", n, N, np.dtype(dtype).itemsize)["lanes"]
    outs = []
    rest = 2
    while rest <= lanes:
        x = torch.full_like(dev[0], float("nan"))
        h.solve("crpcr", *dev, x, n, N, trid.make_opts(rest=rest))
        torch.cuda.synchronize()
        print(f"rest={rest}", flush=True)
        outs.append(x.cpu().numpy())
        check(oracle, outs[-1], a, b, c, d, n, dtype)
        rest *= 2
    h.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_multiwave_shapes_and_io(oracle, trid, dtype):
    # the same system split over 2..16 waves, every I/O mode
    import torch
    n, N = 2048, 33
    a, b, c, d = oracle.generate("weak", 8, n, N, dtype)
    dev = [torch.from_numpy(v).cuda() for v in (a, b, c, d)]
    h = trid.Handle()
    for M in (2, 4, 8, 16):
        # FP64 at 16 unknowns per lane runs 512-thread workgroups (<= 8 waves)
        if n // M > (1024 if dtype == np.float32 or M <= 8 else 512):
            continue
        for io in (1, 2, 3, 4):
            x = torch.full_like(dev[0], float("nan"))
            h.solve("crpcr", *dev, x, n, N,
                    trid.make_opts(rest=32, lanes=n // M, per_lane=M, io_mode=io))
            torch.cuda.synchronize()
            check(oracle, x.cpu().numpy(), a, b, c, d, n, dtype, TOL[dtype] * 10)
    # FP64 n = 8192 at 16 unknowns per lane (8 waves), every I/O mode
    if dtype == np.float64:
        a, b, c, d = oracle.generate("dominant", 10, 8192, 5, dtype)
        dev8 = [torch.from_numpy(v).cuda() for v in (a, b, c, d)]
        for io in (1, 2, 3, 4):
            x = torch.full_like(dev8[0], float("nan"))
            h.solve("crpcr", *dev8, x, 8192, 5,
                    trid.make_opts(rest=64, lanes=512, per_lane=16, io_mode=io))
            torch.cuda.synchronize()
            check(oracle, x.cpu().numpy(), a, b, c, d, 8192, dtype)
    # FP64 n=1024 over two waves
    if dtype == np.float64:
        a, b, c, d = oracle.generate("dominant", 9, 1024, N, dtype)
        dev = [torch.from_numpy(v).cuda() for v in (a, b, c, d)]
        x = torch.full_like(dev[0], float("nan"))
        h.solve("crpcr", *dev, x, 1024, N, trid.make_opts(rest=64, lanes=128, per_lane=8))
        torch.cuda.synchronize()
        check(oracle, x.cpu().numpy(), a, b, c, d, 1024, dtype)
    h.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("kind", ["dominant", "weak"])
def test_8192_persistent_form(oracle, trid, kind, dtype):
    # 8192-unknown dense batches with the LDS-staged I/O modes run persistent
    # workgroups (one per CU in FP64, two in FP32), each walking 2 to 5
    # systems here and copying the next one's a, c into LDS while it solves:
    # every system against the oracle, and bit-identical to the per-lane I/O
    # form (one workgroup per system), for reduced systems in one wave
    # (rest = 8) and across the workgroup (64, 512)
    import torch
    n = 8192
    N = 5 * torch.cuda.get_device_properties(0).multi_processor_count + 5
    a, b, c, d = oracle.generate(kind, 41, n, N, dtype)
    dev = [torch.from_numpy(v).cuda() for v in (a, b, c, d)]
    h = trid.Handle()
    for rest in (8, 64, 512):
        outs = []
        for io in (1, 3, 4):
            x = torch.full_like(dev[0], float("nan"))
            h.solve("crpcr", *dev, x, n, N,
                    trid.make_opts(rest=rest, lanes=512, per_lane=16, io_mode=io))
            torch.cuda.synchronize()
            outs.append(x.cpu().numpy())
        print(f"{np.dtype(dtype).name} 8192 x {N} {kind} rest={rest}", flush=True)
        check(oracle, outs[2], a, b, c, d, n, dtype, TOL[dtype] * (10 if kind == "weak" else 1))
        for o in outs[1:]:
            assert np.array_equal(o, outs[0]), rest
    h.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("kind", ["dominant", "weak"])
def test_4096x8192_full_size(oracle, trid, kind, dtype):
    """4,096 systems x 8,192 unknowns (one system per workgroup, several
    waves each), strongly and weakly (b = |a|+|c|+U[0.01,0.1]) dominant,
    whole batch against FP64 Thomas; CR-PCR and CR."""
    import torch
    n, N = 8192, 4096
    k = 0 if kind == "dominant" else 1
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    dev = [torch.empty(n * N, dtype=tdt, device="cuda") for _ in range(4)]
    trid.generate(k, 2, *dev, n, N)
    a, b, c, d = oracle.generate(kind, 2, n, N, dtype)
    tol = TOL[dtype] * (10 if kind == "weak" else 1)
    h = trid.Handle()
    for method in ("crpcr", "cr"):
        x = torch.full_like(dev[0], float("nan"))
        h.solve(method, *dev, x, n, N)
        torch.cuda.synchronize()
        worst = check(oracle, x.cpu().numpy(), a, b, c, d, n, dtype, tol)
        print(f"4096x8192 {kind} {np.dtype(dtype).name} {method}: max rel err {worst:.3e}",
              flush=True)
    h.close()
    del dev
    torch.cuda.empty_cache()
