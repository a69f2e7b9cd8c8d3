"""GPU parity of the interleaved-batch solve (trid_solve_interleaved_*: row i
of system j at i * batchStride + j, rocSPARSE's gtsv_interleaved_batch
layout) against the FP64 Thomas oracle: the same systems are generated
system-major, transposed into the interleaved layout, solved, and
transposed back. Tolerances as in test_gpu_parity."""
import numpy as np
import pytest

from test_gpu_parity import TOL, check

pytestmark = pytest.mark.gpu


def _interleave(v, n, N, stride):
    out = np.zeros(n * stride, dtype=v.dtype)
    out.reshape(n, stride)[:, :N] = v.reshape(N, n).T
    return out


def _deinterleave(v, n, N, stride):
    return np.ascontiguousarray(v.reshape(n, stride)[:, :N].T).reshape(-1)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("kind", ["dominant", "weak"])
def test_interleaved_matches_oracle(oracle, trid, dtype, kind):
    import torch
    h = trid.Handle()
    # (rows, systems, stride): odd sizes, stride above the system count,
    # one system, two rows, chains whose halves differ by one row
    # (the first, (4096, 64, 64), (301, 4096, 4128) and (2049, 96, 96) run the
    # staged form: system counts in whole groups of 32, 256 rows and more)
    for n, N, stride in ((1024, 4096, 4096), (1000, 3000, 3100), (2, 5, 5), (3, 1, 7),
                         (257, 700, 700), (4096, 64, 64), (301, 4096, 4128), (2049, 96, 96)):
        a, b, c, d = oracle.generate(kind, 17 + n, n, N, dtype)
        dev = [torch.from_numpy(_interleave(v, n, N, stride)).cuda() for v in (a, b, c, d)]
        x = torch.full_like(dev[0], float("nan"))
        print(f"{np.dtype(dtype).name} {kind} n={n} N={N} stride={stride}", flush=True)
        h.solve_interleaved(*dev, x, n, N, stride)
        torch.cuda.synchronize()
        xs = _deinterleave(x.cpu().numpy(), n, N, stride)
        tol = TOL[dtype] * (10 if kind == "weak" else 1)
        check(oracle, xs, a, b, c, d, n, dtype, tol)
        # padding columns (j >= N) are not written
        if stride > N:
            pad = x.cpu().numpy().reshape(n, stride)[:, N:]
            assert np.all(np.isnan(pad))
    h.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_interleaved_in_place(oracle, trid, dtype):
    import torch
    n, N = 512, 2048
    a, b, c, d = oracle.generate("dominant", 5, n, N, dtype)
    dev = [torch.from_numpy(_interleave(v, n, N, N)).cuda() for v in (a, b, c, d)]
    h = trid.Handle()
    x_sep = torch.empty_like(dev[0])
    h.solve_interleaved(*dev, x_sep, n, N)
    h.solve_interleaved(*dev, dev[3], n, N)     # x = d
    torch.cuda.synchronize()
    assert torch.equal(x_sep, dev[3])
    check(oracle, _deinterleave(x_sep.cpu().numpy(), n, N, N), a, b, c, d, n, dtype)
    h.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_interleaved_register_form(oracle, trid, dtype):
    """256 < n <= 1024 (FP32) / 128 < n <= 1024 (FP64) runs one wave per
    system on the register CR-PCR core (16, 8 or 4 unknowns per lane; the
    systems gathered through LDS), shorter FP32 systems the lane-per-system
    elimination; both are checked here: partial
    groups of systems, strides that break 16-byte vectors, an odd number of
    rows, unaligned arrays, x = d in place, and the whole 65,536 x 1,024
    batch."""
    import torch
    h = trid.Handle()
    cases = ((1024, 17, 20, 0), (513, 100, 101, 0), (1000, 3000, 3100, 0), (1024, 4096, 4096, 1),
             (777, 250, 256, 0), (1024, 65536, 65536, 0),
             (512, 300, 300, 0), (300, 1000, 1003, 0), (257, 64, 64, 1), (256, 4096, 4096, 0),
             (200, 17, 20, 0), (129, 100, 101, 0))
    for n, N, stride, offset in cases:
        a, b, c, d = oracle.generate("dominant", 40 + n, n, N, dtype)
        dev = []
        for v in (a, b, c, d):
            t = torch.empty(n * stride + offset, dtype=torch.from_numpy(v[:1]).dtype,
                            device="cuda")[offset:]
            t.copy_(torch.from_numpy(_interleave(v, n, N, stride)))
            dev.append(t)
        x = torch.full_like(dev[0], float("nan"))
        print(f"{np.dtype(dtype).name} register form n={n} N={N} stride={stride} "
              f"offset={offset}", flush=True)
        h.solve_interleaved(*dev, x, n, N, stride)
        torch.cuda.synchronize()
        xs = _deinterleave(x.cpu().numpy(), n, N, stride)
        check(oracle, xs, a, b, c, d, n, dtype)
        if stride > N:
            assert np.all(np.isnan(x.cpu().numpy().reshape(n, stride)[:, N:]))
        # in place: x = d gives the same bits
        h.solve_interleaved(*dev, dev[3], n, N, stride)
        torch.cuda.synchronize()
        assert torch.equal(dev[3].view(n, stride)[:, :N], x.view(n, stride)[:, :N])
    h.close()
