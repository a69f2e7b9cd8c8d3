"""GPU parity of the partial-pivoting solve (trid_solve_pivoting_*,
trid_gtsv_*: LAPACK ?gtsv semantics, the algorithm of the reference's MKL
baseline, tridiagonalMKL.cpp:5-26) against the oracle's ?gtsv restatement,
which test_oracle_gtsv pins bit for bit to scipy's LAPACK.

The kernel performs the same operations in the same order with IEEE
division and no contraction, so the bar is bit equality of x (NaN where the
system is singular) and of INFO, in FP32 and FP64, for both the system-major
and the interleaved layout."""
import numpy as np
import pytest

from test_oracle_gtsv import integer_systems, same_bits

pytestmark = pytest.mark.gpu


def _dev(torch, arrs, offset=0):
    """Device copies; offset > 0 shifts the data by that many elements so
    the pointers are not 16-byte aligned (the kernel's scalar-load form)."""
    out = []
    for v in arrs:
        t = torch.empty(v.size + offset, dtype=torch.from_numpy(v[:1]).dtype, device="cuda")
        t = t[offset:]
        t.copy_(torch.from_numpy(v))
        out.append(t)
    return out


def _solve(trid, torch, h, a, b, c, d, n, offset=0, in_place=False):
    N = a.size // n
    dev = _dev(torch, (a, b, c, d), offset)
    x = dev[3] if in_place else torch.full_like(dev[0], 7.0)
    info = torch.full((max(N, 1),), -5, dtype=torch.int32, device="cuda")
    h.solve_pivoting(*dev, x, n, N, info=info)
    torch.cuda.synchronize()
    return x.cpu().numpy(), info.cpu().numpy()[:N]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("kind", ["paired", "general"])
def test_pivoting_bitwise_equals_oracle(oracle, trid, dtype, kind):
    import torch
    h = trid.Handle()
    # (rows, systems): one row, partial chunks, odd sizes, partial waves,
    # systems longer than one chunk of rows many times over
    for n, N in ((1, 5), (2, 64), (3, 65), (5, 1), (16, 64), (17, 100), (64, 300),
                 (100, 127), (1000, 256), (1024, 1024), (4097, 70)):
        a, b, c, d = oracle.generate(kind, 31 + n, n, N, dtype)
        xr, infor = oracle.gtsv(a, b, c, d, n)
        for offset in ((0, 1) if n % 4 == 0 else (0,)):
            print(f"{np.dtype(dtype).name} {kind} n={n} N={N} offset={offset}", flush=True)
            x, info = _solve(trid, torch, h, a, b, c, d, n, offset)
            assert np.array_equal(info, infor), (n, N, offset)
            assert same_bits(x, xr), (n, N, offset)
    h.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_pivoting_ties_and_singular_systems(oracle, trid, dtype):
    import torch
    h = trid.Handle()
    for n in (1, 2, 3, 4, 6, 33):
        a, b, c, d = integer_systems(100 + n, n, 600, dtype)
        xr, infor = oracle.gtsv(a, b, c, d, n)
        print(f"{np.dtype(dtype).name} integer n={n}", flush=True)
        x, info = _solve(trid, torch, h, a, b, c, d, n)
        assert np.array_equal(info, infor), n
        assert same_bits(x, xr), n
        assert (infor != 0).any() and (infor == 0).any()
    h.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_pivoting_in_place_and_without_info(oracle, trid, dtype):
    import torch
    n, N = 1000, 500
    a, b, c, d = oracle.generate("paired", 8, n, N, dtype)
    xr, _ = oracle.gtsv(a, b, c, d, n)
    h = trid.Handle()
    x, _ = _solve(trid, torch, h, a, b, c, d, n, in_place=True)
    assert same_bits(x, xr)
    dev = _dev(torch, (a, b, c, d))
    x2 = torch.empty_like(dev[0])
    h.solve_pivoting(*dev, x2, n, N)           # info = NULL
    torch.cuda.synchronize()
    assert same_bits(x2.cpu().numpy(), xr)
    h.close()


def _interleave(v, n, N, stride):
    out = np.full(n * stride, np.nan, dtype=v.dtype)
    out.reshape(n, stride)[:, :N] = v.reshape(N, n).T
    return out


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_pivoting_interleaved_bitwise(oracle, trid, dtype):
    import torch
    h = trid.Handle()
    for n, N, stride in ((1024, 4096, 4096), (1000, 3000, 3100), (2, 5, 5), (1, 3, 7),
                         (257, 700, 701), (33, 64, 64)):
        a, b, c, d = oracle.generate("paired", 3 + n, n, N, dtype)
        xr, infor = oracle.gtsv(a, b, c, d, n)
        dev = [torch.from_numpy(_interleave(v, n, N, stride)).cuda() for v in (a, b, c, d)]
        x = torch.full_like(dev[0], 7.0)
        info = torch.full((N,), -5, dtype=torch.int32, device="cuda")
        print(f"{np.dtype(dtype).name} interleaved n={n} N={N} stride={stride}", flush=True)
        h.solve_pivoting(*dev, x, n, N, info=info, batchStride=stride)
        torch.cuda.synchronize()
        xs = x.cpu().numpy().reshape(n, stride)
        assert same_bits(np.ascontiguousarray(xs[:, :N].T).reshape(-1), xr), (n, N)
        assert np.all(xs[:, N:] == 7.0)          # padding columns untouched
        assert np.array_equal(info.cpu().numpy(), infor)
    h.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gtsv_host_dropin(oracle, trid, dtype):
    n, N = 300, 129
    a, b, c, d = oracle.generate("general", 77, n, N, dtype)
    keep = [v.copy() for v in (a, b, c, d)]
    x = np.empty_like(a)
    trid.tridiagonalMKL(a, b, c, d, x, n, N)
    xr, _ = oracle.gtsv(a, b, c, d, n)
    assert same_bits(x, xr)
    for u, v in zip((a, b, c, d), keep):        # inputs are not modified
        assert np.array_equal(u, v)


def test_pivoting_65536x1024_bitwise(oracle, trid):
    """65,536 systems of 1,024 unknowns, FP32, the "paired" kind (pivoting
    needed): every element equals the oracle's sgtsv restatement."""
    import torch
    n, N = 1024, 65536
    dev = [torch.empty(n * N, dtype=torch.float32, device="cuda") for _ in range(4)]
    trid.generate(3, 2024, *dev, n, N)
    x = torch.empty_like(dev[0])
    info = torch.empty(N, dtype=torch.int32, device="cuda")
    h = trid.Handle()
    h.solve_pivoting(*dev, x, n, N, info=info)
    torch.cuda.synchronize()
    a, b, c, d = oracle.generate("paired", 2024, n, N, np.float32)
    xr, infor = oracle.gtsv(a, b, c, d, n)
    assert np.array_equal(info.cpu().numpy(), infor)
    assert same_bits(x.cpu().numpy(), xr)
    h.close()
