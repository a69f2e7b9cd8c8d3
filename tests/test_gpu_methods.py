"""GPU parity of the LDS-resident CR / CR-PCR kernels (crNBC, crpcrNBC,
crNBC32 and the LDS fallback) and of PCR + Thomas (pcrGaussWrapper), against
the FP64 Thomas oracle. Tolerances as in test_gpu_parity: FP32 <= 1e-4,
FP64 <= 1e-11 (max relative infinity-norm error per system)."""
import numpy as np
import pytest

from test_gpu_parity import TOL, check, solve_host

pytestmark = pytest.mark.gpu


def _sizes(dtype):
    return [2, 8, 17, 64, 100, 512, 1024, 3000, 4096] + (
        [8192] if dtype == np.float32 else [])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("method", ["crnbc", "crnbc32", "crpcrnbc"])
def test_nbc_methods(oracle, trid, method, dtype):
    for n in _sizes(dtype):
        n_pad = 1 << (n - 1).bit_length()
        if method == "crpcrnbc" and n_pad < 4:
            continue
        N = 9
        a, b, c, d = oracle.generate("dominant", 300 + n, n, N, dtype)
        print(f"{method} {np.dtype(dtype).name} n={n}", flush=True)
        x = solve_host(trid, method, a, b, c, d, n, N, min(64, n_pad // 2))
        check(oracle, x, a, b, c, d, n, dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_crpcrnbc_rest_sweep(oracle, trid, dtype):
    # rest <= 64 finishes in one wave; rest > 64 runs PCR levels in LDS
    n, N = 2048, 16
    a, b, c, d = oracle.generate("weak", 4, n, N, dtype)
    rest = 2
    while rest <= n // 2:
        print(f"rest={rest}", flush=True)
        x = solve_host(trid, "crpcrnbc", a, b, c, d, n, N, rest)
        check(oracle, x, a, b, c, d, n, dtype, TOL[dtype] * 10)
        rest *= 2


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_lds_pad_is_layout_only(oracle, trid, dtype):
    # the pad map moves addresses, not arithmetic: bit-identical results
    import torch
    n, N = 1024, 64
    a, b, c, d = oracle.generate("dominant", 6, n, N, dtype)
    dev = [torch.from_numpy(v).cuda() for v in (a, b, c, d)]
    h = trid.Handle()
    outs = []
    for pad in (-1, 16, 32, 64, 128):
        x = torch.full_like(dev[0], float("nan"))
        h.solve("crnbc", *dev, x, n, N, trid.make_opts(lds_pad=pad))
        torch.cuda.synchronize()
        outs.append(x.cpu().numpy())
    h.close()
    check(oracle, outs[0], a, b, c, d, n, dtype)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_pcr_long_systems_use_lds_fallback(oracle, trid, dtype):
    for n in (2048, 4096):
        q = trid.plan_query("pcr", n, 4, np.dtype(dtype).itemsize)
        assert q["family"] == 1
        a, b, c, d = oracle.generate("dominant", 12, n, 4, dtype)
        x = solve_host(trid, "pcr", a, b, c, d, n, 4)
        check(oracle, x, a, b, c, d, n, dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_pcrgauss_every_depth(oracle, trid, dtype):
    # depth log2(G) makes the split exact: every G is within tolerance, also
    # on the weakly dominant variant where a shallower depth is not
    for n in (16, 64, 1000, 1024, 4096):
        n_pad = 1 << (n - 1).bit_length()
        for kind in ("dominant", "weak"):
            a, b, c, d = oracle.generate(kind, 40 + n, n, 7, dtype)
            G = 1
            while G <= n_pad // 2:
                x = solve_host(trid, "pcrgauss", a, b, c, d, n, 7, G)
                tol = TOL[dtype] * (10 if kind == "weak" else 1)
                check(oracle, x, a, b, c, d, n, dtype, tol)
                G *= 4 if n >= 1024 else 2


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [4096, 5000, 1100])
def test_pcrgauss_two_pass_form(oracle, trid, dtype, n):
    # n = 5000, 1100: subsystems of unequal length (1100 with G = 512 gives
    # subsystems of 3 and 2 rows; both halves of the two-ended elimination)
    import torch
    N = 6
    a, b, c, d = oracle.generate("dominant", 77, n, N, dtype)
    dev = [torch.from_numpy(v).cuda() for v in (a, b, c, d)]
    h = trid.Handle()
    for G in (1, 2, 16, 64, 256, 512):
        for flags in (0, 1):
            x = torch.full_like(dev[0], float("nan"))
            h.solve("pcrgauss", *dev, x, n, N, trid.make_opts(gaussian=G, flags=flags))
            torch.cuda.synchronize()
            print(f"n={n} G={G} two_pass={flags}", flush=True)
            check(oracle, x.cpu().numpy(), a, b, c, d, n, dtype)
    h.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_pcrgauss_staged_long_chains(oracle, trid, dtype):
    # subsystems of ~1000 rows and more (the staged second pass), including
    # n not a multiple of G (subsystems of unequal length) and several systems
    import torch
    h = trid.Handle()
    for n, N, G in ((1 << 16, 3, 64), ((1 << 16) + 3, 2, 64), (40000, 2, 32), (1 << 17, 2, 128)):
        a, b, c, d = oracle.generate("dominant", n + G, n, N, dtype)
        dev = [torch.from_numpy(v).cuda() for v in (a, b, c, d)]
        x = torch.full_like(dev[0], float("nan"))
        h.solve("pcrgauss", *dev, x, n, N, trid.make_opts(gaussian=G, flags=1))
        torch.cuda.synchronize()
        print(f"n={n} N={N} G={G}", flush=True)
        check(oracle, x.cpu().numpy(), a, b, c, d, n, dtype)
    h.close()


def test_pcrgauss_million_unknowns(oracle, trid):
    # the long-system form (global PCR pass + strided Thomas) at n = 2^20
    import torch
    n, N = 1 << 20, 2
    dev = [torch.empty(n * N, dtype=torch.float32, device="cuda") for _ in range(4)]
    trid.generate(0, 5, *dev, n, N)
    a, b, c, d = oracle.generate("dominant", 5, n, N, np.float32)
    h = trid.Handle()
    for G in (4, 64, 256, 1024):
        x = torch.full_like(dev[0], float("nan"))
        h.solve("pcrgauss", *dev, x, n, N, trid.make_opts(gaussian=G))
        torch.cuda.synchronize()
        print(f"n=2^20 G={G}", flush=True)
        check(oracle, x.cpu().numpy(), a, b, c, d, n, np.float32)
    h.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_long_batch_full_size(oracle, trid, dtype):
    # 4,096 systems x 8,192 unknowns, whole batch against the oracle
    import torch
    n, N = 8192, 4096
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    dev = [torch.empty(n * N, dtype=tdt, device="cuda") for _ in range(4)]
    trid.generate(0, 1, *dev, n, N)
    x = torch.empty_like(dev[0])
    h = trid.Handle()
    h.solve("crpcr", *dev, x, n, N, trid.make_opts(rest=64))
    torch.cuda.synchronize()
    a, b, c, d = oracle.generate("dominant", 1, n, N, dtype)
    worst = check(oracle, x.cpu().numpy(), a, b, c, d, n, dtype)
    print(f"4096x8192 {np.dtype(dtype).name}: {worst:.3e}", flush=True)
    h.close()


def test_fp64_long_systems_batch_independent(oracle, trid):
    """FP64 systems of 8,192 unknowns (16 per lane, 8 waves): every system of
    a 1,000-system batch equals its solve in a 100-system batch bit for bit,
    and the oracle within tolerance; both reduced-system paths (rest <= 8
    waves, rest > 8) are covered."""
    import torch
    n, N, Nsmall = 8192, 1000, 100
    a, b, c, d = oracle.generate("dominant", 33, n, N, np.float64)
    dev = [torch.from_numpy(v).cuda() for v in (a, b, c, d)]
    h = trid.Handle()
    for rest in (2, 8, 64, 256):
        o = trid.make_opts(rest=rest)
        x = torch.full_like(dev[0], float("nan"))
        print(f"fp64 8192 x {N} rest={rest}", flush=True)
        h.solve("crpcr", *dev, x, n, N, o)
        xs = torch.full((n * Nsmall,), float("nan"), dtype=torch.float64, device="cuda")
        h.solve("crpcr", *(v[: n * Nsmall] for v in dev), xs, n, Nsmall, o)
        torch.cuda.synchronize()
        xb = x.cpu().numpy()
        assert np.array_equal(xs.cpu().numpy(), xb[: n * Nsmall]), rest
        check(oracle, xb, a, b, c, d, n, np.float64)
    h.close()
