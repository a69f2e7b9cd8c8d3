"""GPU tests of the request server: mixed-size requests from several client
threads against the FP64 Thomas oracle (FP64-level residuals), each request
bit-identical alone and in a busy mix, operator requests, and ADI requests
on a resident grid against the exact mode amplification."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _request(oracle, rng, seed, max_systems=64, lo=32, hi=16384):
    import tridiagonalsolvers_amd.ragged as R
    k = int(rng.integers(1, max_systems + 1))
    sizes = R.log_uniform_sizes(k, lo, hi, seed=seed)
    offs = R.offsets_of(sizes)
    a, b, c, d = oracle.generate("dominant", seed, 0, 0, np.float64, offsets=offs)
    return offs, (a, b, c, d)


def _check(oracle, offs, arrs, x):
    xref = oracle.thomas(*arrs, offsets=offs)
    rel, _, fails = oracle.compare(x, xref, offsets=offs)
    res = oracle.residual(*arrs, x, offsets=offs)
    assert fails == 0 and float(np.max(rel)) <= 1e-11 and float(np.max(res)) <= 1e-12


def test_server_mixed_requests_from_threads(oracle, trid):
    import tridiagonalsolvers_amd.server as S
    srv = S.Server(hbm_budget_bytes=2 << 30, slots=3, max_batch_unknowns=1 << 23)
    rng = np.random.default_rng(0)
    reqs = [_request(oracle, rng, 100 + i) for i in range(48)]
    xs = [np.full(int(o[-1]), np.nan) for o, _ in reqs]
    tickets = [None] * len(reqs)

    def client(ids):
        for i in ids:
            offs, (a, b, c, d) = reqs[i]
            tickets[i] = srv.submit(d, xs[i], offs, a, b, c)

    ths = [threading.Thread(target=client, args=(range(k, len(reqs), 4),)) for k in range(4)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    results = [srv.wait(t) for t in tickets]
    st = srv.stats()
    print(st, flush=True)
    assert all(r["state"] == "done" and r["status"] == 0 for r in results)
    assert st["peak_device_bytes"] <= st["budget_bytes"]
    for (offs, arrs), x in zip(reqs, xs):
        _check(oracle, offs, arrs, x)
    srv.close()


def test_server_alone_equals_in_the_mix(oracle, trid):
    import tridiagonalsolvers_amd.server as S
    srv = S.Server(hbm_budget_bytes=2 << 30, slots=2)
    rng = np.random.default_rng(7)
    target = _request(oracle, rng, 999, max_systems=40)
    offs, (a, b, c, d) = target
    x_alone = np.full(int(offs[-1]), np.nan)
    assert srv.wait(srv.submit(d, x_alone, offs, a, b, c))["state"] == "done"
    others = [_request(oracle, rng, 500 + i) for i in range(24)]
    outs = [np.full(int(o[-1]), np.nan) for o, _ in others]
    x_mix = np.full(int(offs[-1]), np.nan)
    tk = []
    for i, ((oo, (aa, bb, cc, dd)), xo) in enumerate(zip(others, outs)):
        tk.append(srv.submit(dd, xo, oo, aa, bb, cc))
        if i == 11:
            t_target = srv.submit(d, x_mix, offs, a, b, c)
    res = srv.wait(t_target)
    for t in tk:
        srv.wait(t)
    print(res, flush=True)
    assert np.array_equal(x_alone, x_mix)
    _check(oracle, offs, (a, b, c, d), x_mix)
    srv.close()


def test_server_operator_requests(oracle, trid):
    import tridiagonalsolvers_amd.ragged as R
    import tridiagonalsolvers_amd.server as S
    srv = S.Server(hbm_budget_bytes=1 << 30)
    sizes = np.array([5000, 32, 16384, 700], dtype=np.int64)
    offs = R.offsets_of(sizes)
    a, b, c, _ = oracle.generate("dominant", 3, 0, 0, np.float64, offsets=offs)
    op = srv.register_operator(sizes, a, b, c)
    for seed in range(3):
        d = np.random.default_rng(seed).uniform(-1, 1, int(offs[-1]))
        x = np.full_like(d, np.nan)
        assert srv.wait(srv.submit(d, x, op=op))["state"] == "done"
        _check(oracle, offs, (a, b, c, d), x)
    assert srv.release_operator(op) == 0
    srv.close()


def test_server_adi_grid(oracle, trid):
    import tridiagonalsolvers_amd.server as S
    srv = S.Server(hbm_budget_bytes=1 << 30)
    n = 512
    i = np.arange(1, n + 1)
    u0 = np.outer(np.sin(3 * np.pi * i / (n + 1)), np.sin(2 * np.pi * i / (n + 1)))
    g = srv.create_adi(u0)
    rs = [0.4, 7.5, 1.0]
    tk = [srv.submit_adi(g, r) for r in rs]
    for t in tk:
        assert srv.wait(t)["state"] == "done"
    u = srv.read_adi(g, n)
    f = np.prod([oracle.adi_mode_factor(n, 3, 2, r) for r in rs])
    assert np.max(np.abs(u - f * u0)) <= 1e-12
    assert srv.release_adi(g) == 0
    srv.close()


def test_server_host_buffers_match_pageable(oracle, trid):
    # requests whose arrays live in the server's page-locked memory (DMA,
    # no staging copy) give the same bits as pageable ones, also mixed in
    # one batch
    import tridiagonalsolvers_amd.server as S
    srv = S.Server(hbm_budget_bytes=2 << 30, slots=2, batch_window_us=2000)
    rng = np.random.default_rng(11)
    reqs = [_request(oracle, rng, 500 + i) for i in range(12)]
    xs_page = [np.full(int(o[-1]), np.nan) for o, _ in reqs]
    for i, (offs, (a, b, c, d)) in enumerate(reqs):
        srv.wait(srv.submit(d, xs_page[i], offs, a, b, c))
    pinned, tickets = [], []
    for i, (offs, arrs) in enumerate(reqs):
        n = int(offs[-1])
        pa = [srv.host_array(n) for _ in range(5)]
        for dst, src in zip(pa, arrs):
            dst[:] = src
        pa[4][:] = np.nan
        pinned.append(pa)
        if i % 3 == 2:   # some requests keep pageable inputs in the same batches
            a, b, c, d = arrs
        else:
            a, b, c, d = pa[:4]
        tickets.append(srv.submit(d, pa[4], offs, a, b, c))
    for t in tickets:
        assert srv.wait(t)["state"] == "done"
    for i, (offs, arrs) in enumerate(reqs):
        assert np.array_equal(pinned[i][4], xs_page[i]), i
        _check(oracle, offs, arrs, pinned[i][4])
    srv.close()
