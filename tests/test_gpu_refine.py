"""GPU tests of the FP32-solve / FP64-refinement path (trid_solve_refined_f64):
relative residual ||d - T x||_inf / ||d||_inf <= 1e-12 after two corrections,
agreement with the FP64 Thomas oracle to 1e-11, and FP32-level error with no
correction (the refinement is doing the work)."""
import numpy as np
import pytest

from test_gpu_parity import check

pytestmark = pytest.mark.gpu


def _run(trid, oracle, kind, n, N, iters, method="crpcr", seed=3):
    import torch
    a, b, c, d = oracle.generate(kind, seed, n, N, np.float64)
    dev = [torch.from_numpy(v).cuda() for v in (a, b, c, d)]
    x = torch.full_like(dev[0], float("nan"))
    h = trid.Handle()
    h.solve_refined(method, *dev, x, n, N, iters=iters)
    torch.cuda.synchronize()
    h.close()
    return (a, b, c, d), x.cpu().numpy()


@pytest.mark.parametrize("n", [2, 16, 100, 512, 1000, 1024, 3000, 8192])
def test_refined_reaches_fp64_accuracy(oracle, trid, n):
    N = 23
    for kind in ("dominant", "weak"):
        (a, b, c, d), x = _run(trid, oracle, kind, n, N, iters=2)
        res = oracle.residual(a, b, c, d, x, n=n)
        print(f"n={n} {kind}: max rel residual {res.max():.2e}", flush=True)
        assert res.max() <= 1e-12
        check(oracle, x, a, b, c, d, n, np.float64, 1e-11 if kind == "dominant" else 1e-10)


def test_refinement_steps_reduce_the_error(oracle, trid):
    n, N = 1024, 64
    worst = []
    for iters in (0, 1, 2):
        (a, b, c, d), x = _run(trid, oracle, "weak", n, N, iters)
        worst.append(oracle.residual(a, b, c, d, x, n=n).max())
    print(worst)
    assert 1e-9 < worst[0] < 1e-4        # plain FP32 solve
    assert worst[1] < worst[0] * 1e-3
    assert worst[2] <= 1e-12


@pytest.mark.parametrize("method", ["cr", "pcr", "crpcr"])
def test_refined_methods(oracle, trid, method):
    n, N = 512, 40
    (a, b, c, d), x = _run(trid, oracle, "dominant", n, N, 2, method=method)
    assert oracle.residual(a, b, c, d, x, n=n).max() <= 1e-12
