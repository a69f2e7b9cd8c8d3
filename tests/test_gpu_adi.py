"""GPU tests of the resident ADI grid (Peaceman-Rachford steps of the 2D heat
equation): sine modes against their exact amplification factor, and random
grids against the oracle's FP64 ADI step (tolerance 1e-12 relative), at
sizes from 37 to 16384 points a side (one wave of the solve per 1,024 rows;
odd sizes take the element-wise path)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(n, u0, rs):
    import torch
    from tridiagonalsolvers_amd.adi import AdiGrid
    u = torch.from_numpy(np.ascontiguousarray(u0, dtype=np.float64)).cuda()
    g = AdiGrid(n, u)
    for r in rs:
        g.step(r)
    torch.cuda.synchronize()
    out = u.cpu().numpy().reshape(n, n)
    g.close()
    return out


@pytest.mark.parametrize("n", [64, 100, 1024, 2048, 10000, 16384])
def test_adi_sine_modes(oracle, trid, n):
    i = np.arange(1, n + 1, dtype=np.int64)

    def mode(k):
        # sin(k pi i / (n+1)) with k i reduced modulo the period 2 (n+1) in
        # integers first: at n = 16384, k = n/2 the unreduced argument is
        # ~2.6e4 and its rounding alone (~3e-12) would exceed the bound
        return np.sin(np.pi * ((k * i) % (2 * (n + 1))) / (n + 1))

    for p, q, r in ((1, 1, 0.1), (5, 3, 1.7), (n // 2, 7, 10.0)):
        u0 = np.outer(mode(p), mode(q))
        u1 = _run(n, u0, [r])
        g = oracle.adi_mode_factor(n, p, q, r)
        err = np.max(np.abs(u1 - g * u0))
        print(f"n={n} p={p} q={q} r={r}: err {err:.2e}", flush=True)
        assert err <= 1e-12


@pytest.mark.parametrize("n", [37, 256, 1024, 4097])
def test_adi_random_grid_matches_oracle(oracle, trid, n):
    rng = np.random.default_rng(n)
    u0 = rng.uniform(-1, 1, (n, n))
    rs = [0.1, 3.3, 10.0]
    u = _run(n, u0, rs)
    ref = u0
    for r in rs:
        ref = oracle.adi_step_np(ref, r)
    err = np.max(np.abs(u - ref)) / np.max(np.abs(ref))
    print(f"n={n}: rel err {err:.2e}")
    assert err <= 1e-12


@pytest.mark.parametrize("n", [37, 256, 2048])
def test_adi_several_steps_in_one_call(oracle, trid, n):
    """One call of several steps keeps u' transposed between steps and
    transposes back only at the end; it must equal the steps one by one."""
    import torch
    from tridiagonalsolvers_amd.adi import AdiGrid
    rng = np.random.default_rng(7 + n)
    u0 = rng.uniform(-1, 1, (n, n))
    u = torch.from_numpy(u0.copy()).cuda()
    g = AdiGrid(n, u)
    g.step(2.5, 4)
    torch.cuda.synchronize()
    got = u.cpu().numpy().reshape(n, n)
    g.close()
    ref = u0
    for _ in range(4):
        ref = oracle.adi_step_np(ref, 2.5)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"n={n}: 4 steps, rel err {err:.2e}", flush=True)
    assert err <= 1e-12
    # a sine mode decays by its amplification factor per step
    i = np.arange(1, n + 1)
    s0 = np.outer(np.sin(3 * np.pi * i / (n + 1)), np.sin(2 * np.pi * i / (n + 1)))
    u1 = _run_steps(n, s0, 0.7, 3)
    f = oracle.adi_mode_factor(n, 3, 2, 0.7)
    assert np.max(np.abs(u1 - f ** 3 * s0)) <= 1e-12


def _run_steps(n, u0, r, steps):
    import torch
    from tridiagonalsolvers_amd.adi import AdiGrid
    u = torch.from_numpy(np.ascontiguousarray(u0, dtype=np.float64)).cuda()
    g = AdiGrid(n, u)
    g.step(r, steps)
    torch.cuda.synchronize()
    out = u.cpu().numpy().reshape(n, n)
    g.close()
    return out
