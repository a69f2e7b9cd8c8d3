"""The served ADI workload on the GPU at its configured sizes: resident ADI
grids of 2048^2 and 4096^2 stepped (r ~ U[0.1, 10]) while Poisson-arriving
solve requests of 1..4,096 systems with sizes log-uniform in [32, 16,384]
(FP32 solve + 2 FP64 refinements) are in flight on the same server.

Checked: every ADI step against the exact amplification of sine modes (both
sizes) and, on a random 2048^2 grid, against the oracle's FP64 ADI step
applied in the same order, to 1e-12; every request bit-identical to the same
request solved alone afterwards, with relative residual <= 1e-12 (long
double, host) and error <= 1e-11 against FP64 Thomas."""
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_served_adi_grids_with_poisson_requests(oracle, trid):
    import torch
    import tridiagonalsolvers_amd.ragged as R
    import tridiagonalsolvers_amd.server as S

    rng = np.random.default_rng(42)
    pool_n = 1 << 24
    max_req = 1 << 23
    pool = [torch.empty(pool_n, dtype=torch.float64, device="cuda") for _ in range(4)]
    trid.generate(0, 99, *pool, pool_n, 1)
    host = [v.cpu().numpy() for v in pool]
    del pool
    reqs = []
    while len(reqs) < 16:
        k = int(rng.integers(1, 4097))
        sizes = R.log_uniform_sizes(k, 32, 16384, seed=int(rng.integers(1 << 30)))
        tot = int(sizes.sum())
        if tot > max_req:
            continue
        start = int(rng.integers(0, pool_n - tot))
        reqs.append((R.offsets_of(sizes), start, tot))
    arrays = [[h[s:s + t] for h in host] for _, s, t in reqs]
    print(f"served: {len(reqs)} requests, {sum(r[2] for r in reqs)} unknowns", flush=True)

    srv = S.Server(hbm_budget_bytes=8 << 30, slots=3, max_batch_unknowns=max_req,
                   batch_window_us=300)
    # resident grids: sine modes at both sizes, a random 2048^2 grid
    modes = {2048: (3, 2), 4096: (5, 7)}
    u0s, grids = {}, {}
    for n, (p, q) in modes.items():
        i = np.arange(1, n + 1)
        u0s[n] = np.outer(np.sin(p * np.pi * i / (n + 1)), np.sin(q * np.pi * i / (n + 1)))
        grids[n] = srv.create_adi(u0s[n])
    u_rand = np.random.default_rng(5).random((2048, 2048))
    g_rand = srv.create_adi(u_rand)

    lam = 150.0            # requests per second
    steps = {2048: [], 4096: [], "rand": []}
    adi_tickets = []
    stop = threading.Event()

    def adi_client():
        r_rng = np.random.default_rng(7)
        while not stop.is_set():
            for key, gid in ((2048, grids[2048]), (4096, grids[4096]), ("rand", g_rand)):
                if key == "rand" and len(steps["rand"]) >= 4:
                    continue
                r = float(r_rng.uniform(0.1, 10.0))
                steps[key].append(r)
                adi_tickets.append(srv.submit_adi(gid, r))
            time.sleep(float(r_rng.exponential(4.0 / lam)))

    xs = [np.full(r[2], np.nan) for r in reqs]
    th = threading.Thread(target=adi_client)
    th.start()
    tickets = []
    gaps = rng.exponential(1.0 / lam, size=len(reqs))
    for i, gap in enumerate(gaps):
        time.sleep(float(gap))
        offs = reqs[i][0]
        a, b, c, d = arrays[i]
        tickets.append(srv.submit(d, xs[i], offs, a, b, c))
    res = [srv.wait(t) for t in tickets]
    stop.set()
    th.join()
    adi_res = [srv.wait(t) for t in adi_tickets]
    assert all(r["state"] == "done" and r["status"] == 0 for r in res)
    assert all(r["state"] == "done" and r["status"] == 0 for r in adi_res)
    lat = sorted(r["latency_us"] / 1e3 for r in res)
    print(f"served: request latency p50 {lat[len(lat) // 2]:.2f} ms, max {lat[-1]:.2f} ms; "
          f"ADI steps {len(adi_res)}", flush=True)
    st = srv.stats()
    assert st["peak_device_bytes"] <= st["budget_bytes"]

    # ADI: sine modes scale by the product of the exact one-step factors
    for n, (p, q) in modes.items():
        assert steps[n], n
        u = srv.read_adi(grids[n], n)
        f = float(np.prod([oracle.adi_mode_factor(n, p, q, r) for r in steps[n]]))
        err = float(np.max(np.abs(u - f * u0s[n])))
        print(f"ADI {n}^2 sine mode: {len(steps[n])} steps, max abs err {err:.2e}", flush=True)
        assert err <= 1e-12
    # ADI: the random grid against the oracle step, in submission order
    ref = u_rand
    for r in steps["rand"]:
        ref = oracle.adi_step_np(ref, r)
    u = srv.read_adi(g_rand, 2048)
    err = float(np.max(np.abs(u - ref)) / np.max(np.abs(ref)))
    print(f"ADI 2048^2 random grid: {len(steps['rand'])} steps, max rel err {err:.2e}",
          flush=True)
    assert err <= 1e-12

    # requests: alone = in the mix, bit for bit; FP64-level accuracy
    worst_rel = worst_res = 0.0
    for i, (offs, _, tot) in enumerate(reqs):
        a, b, c, d = arrays[i]
        x1 = np.full(tot, np.nan)
        assert srv.wait(srv.submit(d, x1, offs, a, b, c))["state"] == "done"
        assert np.array_equal(x1, xs[i]), f"request {i}: differs alone and in the mix"
        xref = oracle.thomas(a, b, c, d, offsets=offs)
        rel, _, fails = oracle.compare(xs[i], xref, offsets=offs)
        rr = oracle.residual(a, b, c, d, xs[i], offsets=offs)
        assert fails == 0
        worst_rel = max(worst_rel, float(np.max(rel)))
        worst_res = max(worst_res, float(np.max(rr)))
    print(f"served requests: max rel err {worst_rel:.2e}, max rel residual {worst_res:.2e}",
          flush=True)
    assert worst_rel <= 1e-11 and worst_res <= 1e-12
    for gid in list(grids.values()) + [g_rand]:
        assert srv.release_adi(gid) == 0
    srv.close()
