"""GPU parity of the strided batch (trid_solve_strided_*: system s at
s * batchStride, the layout of rocSPARSE's gtsv_no_pivot_strided_batch)
against the FP64 Thomas oracle. The systems are generated contiguous and
spread to the stride; the gaps between systems must stay untouched, and
batchStride == systemSize must give the contiguous solve bit for bit."""
import numpy as np
import pytest

from test_gpu_parity import TOL, check

pytestmark = pytest.mark.gpu


def _spread(v, n, N, bs, fill):
    out = np.full(N * bs, fill, dtype=v.dtype)
    out.reshape(N, bs)[:, :n] = v.reshape(N, n)
    return out


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("method", ["crpcr", "cr", "pcr"])
def test_strided_matches_oracle(oracle, trid, dtype, method):
    import torch
    h = trid.Handle()
    # one system per wave (wave-staged I/O), several per wave (per-lane
    # vectors), unaligned systems (element I/O), multi-wave systems, a system
    # size that is not a power of two
    cases = [(1024, 300, 1100), (64, 1000, 80), (64, 333, 65), (1000, 50, 1003)]
    if method != "pcr":
        cases.append((8192, 40, 8200))
        # past the register sizes (the separator form): 16-byte path, element
        # path (odd stride), a partial last chunk
        cases += [(40000, 3, 40960), (40000, 3, 40003), (17000, 4, 20000)]
    for n, N, bs in cases:
        a, b, c, d = oracle.generate("dominant", 50 + n, n, N, dtype)
        dev = [torch.from_numpy(_spread(v, n, N, bs, 7.0)).cuda() for v in (a, b, c, d)]
        x = torch.full_like(dev[0], float("nan"))
        print(f"{np.dtype(dtype).name} {method} n={n} N={N} stride={bs}", flush=True)
        h.solve_strided(method, *dev, x, n, N, bs)
        torch.cuda.synchronize()
        xs = x.cpu().numpy().reshape(N, bs)
        check(oracle, np.ascontiguousarray(xs[:, :n]).reshape(-1), a, b, c, d, n, dtype)
        assert np.all(np.isnan(xs[:, n:]))      # gaps untouched
    h.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_stride_equal_to_size_is_the_contiguous_solve(oracle, trid, dtype):
    import torch
    n, N = 1024, 513
    a, b, c, d = oracle.generate("weak", 8, n, N, dtype)
    dev = [torch.from_numpy(v).cuda() for v in (a, b, c, d)]
    x1 = torch.empty_like(dev[0])
    x2 = torch.empty_like(dev[0])
    h = trid.Handle()
    h.solve("crpcr", *dev, x1, n, N)
    h.solve_strided("crpcr", *dev, x2, n, N, n)
    torch.cuda.synchronize()
    assert torch.equal(x1, x2)
    h.close()
