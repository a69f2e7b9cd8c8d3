"""GPU limits: empty batches are no-ops, sizes beyond the kernels are refused
with TRID_ENOTSUP (never a wrong answer), and batches whose element offsets
pass 2^31 solve correctly (64-bit offsets throughout). The large cases need
about 45 GB and 90 GB of device memory and are skipped on smaller devices."""
import numpy as np
import pytest

from test_gpu_parity import TOL, check

pytestmark = pytest.mark.gpu


def _free_bytes():
    import torch
    free, _ = torch.cuda.mem_get_info()
    return free


def test_empty_batch_is_a_noop(trid):
    import torch
    h = trid.Handle()
    for dt in (torch.float32, torch.float64):
        a, b, c, d = (torch.ones(64, dtype=dt, device="cuda") for _ in range(4))
        x = torch.full_like(a, 7.0)
        for m in ("cr", "pcr", "crpcr", "crnbc", "crpcrnbc", "crnbc32", "pcrgauss"):
            h.solve(m, a, b, c, d, x, 64, 0)
        h.solve_interleaved(a, b, c, d, x, 64, 0)
        if dt == torch.float64:
            h.solve_refined("crpcr", a, b, c, d, x, 64, 0)
        torch.cuda.synchronize()
        assert bool((x == 7.0).all())
    h.close()


def test_sizes_beyond_the_kernels_are_refused(trid):
    import torch
    h = trid.Handle()
    x64 = torch.zeros(32768, dtype=torch.float64, device="cuda")
    x32 = torch.zeros(65536, dtype=torch.float32, device="cuda")
    # PCR beyond the register path (cr and crpcr go on by separators)
    cases = [
        ("pcr", x64, 16384),     # FP64 register path ends at 8192
        ("pcr", x32, 32768),     # FP32 register path ends at 16384
    ]
    for m, v, n in cases:
        with pytest.raises(trid.TridError) as e:
            h.solve(m, v, v, v, v, v, n, 1)
        assert e.value.status == 4, (m, n)  # TRID_ENOTSUP
    # pcrGauss two-pass form: G above 1024 (FP32)
    with pytest.raises(trid.TridError) as e:
        h.solve("pcrgauss", x32, x32, x32, x32, x32, 65536, 1, trid.make_opts(gaussian=2048))
    assert e.value.status == 4
    h.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_pcrgauss_two_pass_beyond_65535_systems(oracle, trid, dtype):
    """The two-pass pcrGauss form runs a one-dimensional grid of tiles x
    systems, so batches of more than 65,535 systems solve (70,000 here)."""
    import torch
    n, N = 64, 70000
    a, b, c, d = oracle.generate("dominant", 12, n, N, dtype)
    dev = [torch.from_numpy(v).cuda() for v in (a, b, c, d)]
    x = torch.full_like(dev[0], float("nan"))
    h = trid.Handle()
    h.solve("pcrgauss", *dev, x, n, N, trid.make_opts(gaussian=4, flags=1))
    torch.cuda.synchronize()
    check(oracle, x.cpu().numpy(), a, b, c, d, n, dtype)
    h.close()


def test_offsets_beyond_2_31_elements(oracle, trid):
    import torch
    n, N = 1024, (1 << 21) + 3          # 2,147,486,720 elements per array
    if _free_bytes() < 5 * n * N * 4 * 1.05:
        pytest.skip("needs about 45 GB of free device memory")
    assert n * N > 2**31
    dev = [torch.empty(n * N, dtype=torch.float32, device="cuda") for _ in range(5)]
    a, b, c, d, x = dev
    trid.generate(0, 123, a, b, c, d, n, N)
    h = trid.Handle()
    h.solve("crpcr", a, b, c, d, x, n, N)
    torch.cuda.synchronize()
    for s in (0, 1, (1 << 21) - 1, N - 2, N - 1):
        lo, hi = s * n, (s + 1) * n
        host = [v[lo:hi].cpu().numpy() for v in (a, b, c, d, x)]
        print(f"system {s}: offset {lo}", flush=True)
        check(oracle, host[4], *host[:4], n, np.float32)
    h.close()
    del dev
    torch.cuda.empty_cache()


def test_ragged_offsets_beyond_2_31_elements(oracle, trid):
    import torch
    import tridiagonalsolvers_amd.ragged as R
    n, N = 8192, 262145                 # 2,147,491,840 unknowns, FP64
    total = n * N
    if _free_bytes() < 5 * total * 8 * 1.05:
        pytest.skip("needs about 90 GB of free device memory")
    assert total > 2**31
    dev = [torch.empty(total, dtype=torch.float64, device="cuda") for _ in range(5)]
    a, b, c, d, x = dev
    trid.generate(0, 321, a, b, c, d, total, 1)
    offs = R.offsets_of([n] * N)
    plan = R.RaggedPlan(offs, table="default-fp64")
    plan.solve(a, b, c, d, x, fp64=True, graph=False)
    torch.cuda.synchronize()
    for s in (0, N - 2, N - 1):
        lo, hi = int(offs[s]), int(offs[s + 1])
        host = [v[lo:hi].cpu().numpy() for v in (a, b, c, d, x)]
        print(f"ragged system {s}: offset {lo}", flush=True)
        check(oracle, host[4], *host[:4], n, np.float64, TOL[np.float64])
    plan.close()
    del dev
    torch.cuda.empty_cache()
