"""GPU tests of the ragged persistent solve (trid_ragged_*) in both modes
(FP32 solve + FP64 refinement; one FP64 solve): every system of a
mixed-size batch against the FP64 Thomas oracle, FP64-level residuals,
bit-identical results alone and in the mix, graph replays equal to the
eager launch, and every autotuner candidate within tolerance."""
import numpy as np
import pytest

from test_gpu_parity import check

pytestmark = pytest.mark.gpu

EDGE_SIZES = [2, 3, 5, 16, 17, 100, 511, 512, 513, 1000, 1024, 1025, 2048, 3000, 4096,
              5000, 8191, 8192, 8193, 12000, 16384]
# the FP64 mode takes systems up to 8192 unknowns
EDGE_SIZES_FP64 = [n for n in EDGE_SIZES if n <= 8192]
MODES = ["refine", "fp64"]


def _table(mode):
    """The shipped autotune table of the mode (tune/ragged_f64_refine.json or
    tune/ragged_f64_fp64.json)."""
    return "default-fp64" if mode == "fp64" else "default"


def _solve(plan, dev, x, mode, **kw):
    if mode == "fp64":
        plan.solve(*dev, x, fp64=True, **kw)
    elif mode == "tol":     # refine to relative residual <= 1e-12, one or two steps
        plan.solve(*dev, x, iters=2, tol=1e-12, **kw)
    else:
        plan.solve(*dev, x, iters=2, **kw)


def _batch(oracle, trid_ragged, sizes, seed=11, kind="dominant"):
    import torch
    offs = trid_ragged.offsets_of(sizes)
    a, b, c, d = oracle.generate(kind, seed, 0, 0, np.float64, offsets=offs)
    dev = [torch.from_numpy(v).cuda() for v in (a, b, c, d)]
    return offs, (a, b, c, d), dev


@pytest.fixture(scope="module")
def rg():
    import tridiagonalsolvers_amd.ragged as R
    return R


def _check_ragged(oracle, offs, host, x, tol):
    a, b, c, d = host
    xref = oracle.thomas(a, b, c, d, offsets=offs)
    rel, rms, fails = oracle.compare(x, xref, offsets=offs)
    res = oracle.residual(a, b, c, d, x, offsets=offs)
    assert fails == 0
    assert float(np.max(rel)) <= tol, float(np.max(rel))
    assert float(np.max(res)) <= 1e-12, float(np.max(res))
    return float(np.max(rel)), float(np.max(res))


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("kind", ["dominant", "weak"])
def test_ragged_matches_oracle(oracle, trid, rg, kind, mode):
    import torch
    edge = EDGE_SIZES if mode == "refine" else EDGE_SIZES_FP64
    sizes = np.concatenate([edge, rg.log_uniform_sizes(1500, 16, 8192, seed=5)])
    offs, host, dev = _batch(oracle, rg, sizes, kind=kind)
    x = torch.full_like(dev[0], float("nan"))
    plan = rg.RaggedPlan(offs, table=_table(mode))
    info = plan.info()
    print(info, flush=True)
    assert info["num_systems"] == len(sizes)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        _solve(plan, dev, x, mode, stream=s)
    s.synchronize()
    worst = _check_ragged(oracle, offs, host, x.cpu().numpy(),
                          1e-11 if kind == "dominant" else 1e-10)
    print(f"{kind} {mode}: max rel err {worst[0]:.2e}, max rel residual {worst[1]:.2e}")
    # the device residual kernel: <= 1e-12 here, and equal to the oracle's
    # where the residual is well above rounding level (perturbed unknowns)
    d_offs = torch.from_numpy(offs).cuda()
    rr = torch.empty(len(sizes), dtype=torch.float64, device="cuda")
    rg.residual(*dev, x, d_offs, len(sizes), rr)
    torch.cuda.synchronize()
    assert float(rr.max()) <= 1e-12
    xp = x.cpu().numpy()
    for s in range(0, len(sizes), 7):
        xp[offs[s] + (offs[s + 1] - offs[s]) // 2] += 1e-5
    rg.residual(*dev, torch.from_numpy(xp).cuda(), d_offs, len(sizes), rr)
    torch.cuda.synchronize()
    ref = oracle.residual(*host, xp, offsets=offs)
    sel = np.arange(0, len(sizes), 7)
    assert np.allclose(rr.cpu().numpy()[sel], ref[sel], rtol=1e-9, atol=0)
    plan.close()


@pytest.mark.parametrize("mode", MODES)
def test_ragged_alone_equals_in_the_mix(oracle, trid, rg, mode):
    import torch
    edge = EDGE_SIZES if mode == "refine" else EDGE_SIZES_FP64
    sizes = np.concatenate([edge, rg.log_uniform_sizes(300, 2, 8192, seed=9)])
    offs, host, dev = _batch(oracle, rg, sizes, seed=4)
    x = torch.full_like(dev[0], float("nan"))
    plan = rg.RaggedPlan(offs, table=_table(mode))
    _solve(plan, dev, x, mode, graph=False)
    torch.cuda.synchronize()
    xm = x.cpu().numpy()
    plan.close()
    for s in list(range(len(edge))) + [40, 77, 150]:
        o, e = int(offs[s]), int(offs[s + 1])
        one = [v[o:e].contiguous() for v in dev]
        xs = torch.full_like(one[0], float("nan"))
        p1 = rg.RaggedPlan(np.array([0, e - o], dtype=np.int64), table=_table(mode))
        _solve(p1, one, xs, mode, graph=False)
        torch.cuda.synchronize()
        p1.close()
        assert np.array_equal(xs.cpu().numpy(), xm[o:e]), (s, e - o)


def test_ragged_refines_to_a_residual_target(oracle, trid, rg):
    """trid_ragged_solve_tol_f64: every system ends at or below the target
    (the oracle's long-double residual), a target no system meets gives the
    fixed-count result bit for bit, a target every FP32 solve meets gives the
    unrefined result bit for bit, and each system decides for itself: alone
    and in the mix its result is the same."""
    import torch
    sizes = np.concatenate([EDGE_SIZES, rg.log_uniform_sizes(1500, 16, 8192, seed=31)])
    for kind in ("dominant", "weak"):
        offs, host, dev = _batch(oracle, rg, sizes, seed=17, kind=kind)
        plan = rg.RaggedPlan(offs)

        def run(**kw):
            x = torch.full_like(dev[0], float("nan"))
            plan.solve(*dev, x, graph=False, **kw)
            torch.cuda.synchronize()
            return x.cpu().numpy()

        for tol in (1e-12, 1e-9):
            x = run(iters=8, tol=tol)
            res = oracle.residual(*host, x, offsets=offs)
            print(f"{kind} tol {tol:g}: max rel residual {float(np.max(res)):.2e}", flush=True)
            assert float(np.max(res)) <= tol, (kind, tol, float(np.max(res)))
        assert np.array_equal(run(iters=2, tol=1e-300), run(iters=2))
        assert np.array_equal(run(iters=2, tol=1.0), run(iters=0))
        xm = run(iters=8, tol=1e-12)
        plan.close()
        for s in list(range(len(EDGE_SIZES))) + [40, 300, 900, 1400]:
            o, e = int(offs[s]), int(offs[s + 1])
            one = [v[o:e].contiguous() for v in dev]
            xs = torch.full_like(one[0], float("nan"))
            p1 = rg.RaggedPlan(np.array([0, e - o], dtype=np.int64))
            p1.solve(*one, xs, iters=8, tol=1e-12, graph=False)
            torch.cuda.synchronize()
            p1.close()
            assert np.array_equal(xs.cpu().numpy(), xm[o:e]), (kind, s, e - o)


def test_ragged_graph_replays_match_eager(oracle, trid, rg):
    import torch
    sizes = rg.log_uniform_sizes(2000, 16, 8192, seed=21)
    offs, host, dev = _batch(oracle, rg, sizes, seed=8)
    plan = rg.RaggedPlan(offs)
    refs = {}
    for mode in MODES:
        x_eager = torch.full_like(dev[0], float("nan"))
        _solve(plan, dev, x_eager, mode, graph=False)
        torch.cuda.synchronize()
        refs[mode] = x_eager.cpu().numpy()
    x = torch.full_like(dev[0], float("nan"))
    s = torch.cuda.Stream()
    # same pointers, alternating modes: each change of mode captures anew
    for rep, mode in enumerate(["refine", "refine", "fp64", "fp64", "refine", "fp64"]):
        x.fill_(float("nan"))
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            _solve(plan, dev, x, mode, stream=s)
        s.synchronize()
        assert np.array_equal(x.cpu().numpy(), refs[mode]), (rep, mode)
        del junk
    plan.close()
    for mode in MODES:
        _check_ragged(oracle, offs, host, refs[mode], 1e-11)


def test_ragged_fp64_mode_limits(oracle, trid, rg):
    import torch
    from tridiagonalsolvers_amd._lib import TridError
    offs, host, dev = _batch(oracle, rg, [100, 16384, 30], seed=3)
    x = torch.zeros_like(dev[0])
    plan = rg.RaggedPlan(offs)
    with pytest.raises(TridError):       # 16384 unknowns: refine mode only
        plan.solve(*dev, x, fp64=True, graph=False)
    plan.close()
    offs, host, dev = _batch(oracle, rg, [100, 8192, 30], seed=3)
    plan = rg.RaggedPlan(offs)
    with pytest.raises(TridError):       # no refinement passes in the FP64 mode
        plan.solve(*dev, x[:8322], iters=2, fp64=True, graph=False)
    plan.close()


@pytest.mark.parametrize("mode", MODES)
def test_every_tuner_candidate_is_accurate(oracle, trid, rg, mode):
    import torch
    from tridiagonalsolvers_amd.autotune import candidates
    for size in (4, 64, 512, 1024, 2048, 4096, 8192, 16384):
        if mode == "fp64" and size > 8192:
            continue
        n = size - (size // 5)   # not a power of two: padded
        n = max(n, 2)
        sizes = [n] * 13         # the last item of a bucket partly filled
        offs, host, dev = _batch(oracle, rg, sizes, seed=size)
        for cand in candidates(size):
            x = torch.full_like(dev[0], float("nan"))
            plan = rg.RaggedPlan(offs, table=[cand])
            _solve(plan, dev, x, mode, graph=False)
            torch.cuda.synchronize()
            plan.close()
            _check_ragged(oracle, offs, host, x.cpu().numpy(), 1e-11)


@pytest.mark.parametrize("mode", MODES + ["tol"])
def test_ragged_100k_full_size(oracle, trid, rg, mode):
    """The configured ragged workload at its full size: 100,000 systems with
    sizes log-uniform in [16, 8192] (the benchmark's batch), each mode with
    its shipped table, every system against FP64 Thomas and its residual."""
    import torch
    sizes = np.sort(rg.log_uniform_sizes(100000, 16, 8192, seed=1))
    offs = rg.offsets_of(sizes)
    total = int(offs[-1])
    print(f"100k ragged {mode}: {total} unknowns", flush=True)
    dev = [torch.empty(total, dtype=torch.float64, device="cuda") for _ in range(4)]
    trid.generate(0, 2024, *dev, total, 1)   # a[first], c[last] of each system are ignored
    x = torch.full_like(dev[0], float("nan"))
    plan = rg.RaggedPlan(offs, table=_table(mode))
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        _solve(plan, dev, x, mode, stream=s)
    s.synchronize()
    plan.close()
    host = [v.cpu().numpy() for v in dev]
    rel, res = _check_ragged(oracle, offs, host, x.cpu().numpy(), 1e-11)
    print(f"100k ragged {mode}: max rel err {rel:.2e}, max rel residual {res:.2e}", flush=True)
    del dev, x
    torch.cuda.empty_cache()
