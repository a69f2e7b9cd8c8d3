"""GPU parity of CR / CR-PCR by separators (systems longer than the register
kernels hold: above 16,384 unknowns in FP32, 8,192 in FP64) against the FP64
Thomas oracle (tridiagonal_gold.h:26-44), with the tolerances of the other
solves: max relative error <= 1e-4 in FP32, <= 1e-11 in FP64.

The form cuts a system into chunks of 1,024 rows, reduces each chunk to its
separator row, solves the separator system and solves each chunk again with
both separators known. The cases cover whole and partial last chunks, both
load paths (16-byte vectors when n is a multiple of 16 bytes, elements
otherwise), every switch size of the chunk solves, x in place of d, and
separator systems long enough to take the same form again."""
import numpy as np
import pytest

from test_gpu_parity import TOL, check

pytestmark = pytest.mark.gpu


def _solve(trid, torch, method, arrs, n, N, opts=None, in_place=False, offset=0):
    dev = []
    for v in arrs:
        t = torch.empty(v.size + offset, dtype=torch.from_numpy(v[:1]).dtype, device="cuda")
        t = t[offset:]
        t.copy_(torch.from_numpy(v))
        dev.append(t)
    x = dev[3] if in_place else torch.full_like(dev[0], float("nan"))
    h = trid.Handle()
    h.solve(method, *dev, x, n, N, opts)
    torch.cuda.synchronize()
    h.close()
    return x.cpu().numpy()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("method", ["cr", "crpcr"])
def test_separator_form_sizes(oracle, trid, method, dtype):
    import torch
    small = 16385 if dtype == np.float32 else 8193
    # (n, N): just past the register sizes (element path, partial last
    # chunk), whole chunks, a partial last chunk on the vector path, odd
    # sizes, one system, several
    for n, N in ((small, 3), (small + 3, 2), (32768, 4), (100000, 2), (65536 + 1023, 3),
                 (262144, 1), (300001, 2)):
        a, b, c, d = oracle.generate("dominant", 7 + n, n, N, dtype)
        print(f"{method} {np.dtype(dtype).name} n={n} N={N}", flush=True)
        x = _solve(trid, torch, method, (a, b, c, d), n, N)
        worst = check(oracle, x, a, b, c, d, n, dtype)
        print(f"  max rel err {worst:.3e}", flush=True)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_separator_form_switch_sizes(oracle, trid, dtype):
    import torch
    n, N = 50000, 3
    a, b, c, d = oracle.generate("dominant", 5, n, N, dtype)
    for rest in (2, 4, 16, 64, 128, 256, 512):
        print(f"{np.dtype(dtype).name} rest={rest}", flush=True)
        x = _solve(trid, torch, "crpcr", (a, b, c, d), n, N, trid.make_opts(rest=rest))
        check(oracle, x, a, b, c, d, n, dtype)
    with pytest.raises(trid.TridError) as e:    # beyond half a chunk
        _solve(trid, torch, "crpcr", (a, b, c, d), n, N, trid.make_opts(rest=1024))
    assert e.value.status == 4


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_separator_form_in_place_unaligned_and_weak(oracle, trid, dtype):
    import torch
    n, N = 70000, 2
    a, b, c, d = oracle.generate("dominant", 21, n, N, dtype)
    x = _solve(trid, torch, "crpcr", (a, b, c, d), n, N, in_place=True)
    check(oracle, x, a, b, c, d, n, dtype)
    x = _solve(trid, torch, "crpcr", (a, b, c, d), n, N, offset=1)   # element path
    check(oracle, x, a, b, c, d, n, dtype)
    a, b, c, d = oracle.generate("weak", 22, n, N, dtype)
    x = _solve(trid, torch, "crpcr", (a, b, c, d), n, N)
    check(oracle, x, a, b, c, d, n, dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_separator_system_takes_the_same_form(oracle, trid, dtype):
    """More than 16,384 (FP32) / 8,192 (FP64) chunks: the separator system
    is itself solved by separators."""
    import torch
    K = (16384 if dtype == np.float32 else 8192) + 1
    n = K * 1024 - 100
    assert trid.plan_query("crpcr", K, 1, np.dtype(dtype).itemsize)["family"] == 3
    a, b, c, d = oracle.generate("dominant", 3, n, 1, dtype)
    print(f"{np.dtype(dtype).name} n={n}", flush=True)
    x = _solve(trid, torch, "crpcr", (a, b, c, d), n, 1)
    worst = check(oracle, x, a, b, c, d, n, dtype)
    print(f"  max rel err {worst:.3e}", flush=True)


def test_64x1048576_full_size(oracle, trid):
    """64 systems of 1,048,576 unknowns, FP32, crpcr: the whole batch."""
    import torch
    n, N = 1 << 20, 64
    dev = [torch.empty(n * N, dtype=torch.float32, device="cuda") for _ in range(4)]
    trid.generate(0, 9, *dev, n, N)
    x = torch.empty_like(dev[0])
    h = trid.Handle()
    h.solve("crpcr", *dev, x, n, N)
    torch.cuda.synchronize()
    h.close()
    a, b, c, d = oracle.generate("dominant", 9, n, N, np.float32)
    worst = check(oracle, x.cpu().numpy(), a, b, c, d, n, np.float32)
    print(f"64x1048576 f32: max rel err {worst:.3e}", flush=True)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_host_dropins_past_the_register_sizes(oracle, trid, dtype):
    """The reference-named host entry points (crpcr_app.cu:5, cr_app.cu:4)
    take the separator form for long systems, with a workspace of their own."""
    n, N = 40000, 3
    a, b, c, d = oracle.generate("dominant", 44, n, N, dtype)
    for name, args in (("crpcr", (64,)), ("cr", ())):
        x = np.full_like(a, np.nan)
        getattr(trid, name)(a, b, c, d, x, n, N, *args)
        check(oracle, x, a, b, c, d, n, dtype)


def test_separator_form_replays_in_a_hip_graph(oracle, trid):
    """Once a first call has sized the handle's workspace, the separator
    form's launches (no allocation, no memset) can be captured: every replay
    equals the eager result bit for bit, with eager work in between."""
    import torch
    n, N = 50000, 4
    a, b, c, d = oracle.generate("dominant", 61, n, N, np.float32)
    dev = [torch.from_numpy(v).cuda() for v in (a, b, c, d)]
    h = trid.Handle()
    x_eager = torch.empty_like(dev[0])
    h.solve("crpcr", *dev, x_eager, n, N)            # sizes the workspace
    torch.cuda.synchronize()
    x = torch.full_like(dev[0], float("nan"))
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        h.solve("crpcr", *dev, x, n, N, stream=s)
    for rep in range(3):
        x.fill_(float("nan"))
        junk = torch.randn(1 << 20, device="cuda").sum()   # eager work between replays
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(x, x_eager), rep
        assert bool(torch.isfinite(junk))
    check(oracle, x.cpu().numpy(), a, b, c, d, n, np.float32)
    h.close()
