"""Benchmark: batched CR-PCR, FP32, 65,536 systems x 1,024 unknowns (default).

One step = one solve of the whole batch (inputs already resident in HBM,
written by the device generator, seed 1, diagonally dominant kind). Prints
ONE JSON line with unknowns solved per second, the HBM roofline of the solve
kernel (algorithmic bytes 5 * n * 4 per system: read a, b, c, d, write x)
and a CPU baseline (the FP64 and FP32 Thomas recurrences of the oracle,
OpenMP over all host cores) timed in the same run, then the served half of
the metric: the request server under Poisson arrivals with resident ADI
grids (p50/p99 latency, served unknowns/s against plain per-size launches,
max relative residual) in a `served` object.

    python bench.py [--steps K] [--warmup W] [--method crpcr] [--n 1024]
                    [--systems 65536] [--rest 64] [--dtype f32]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E, AMD published peak
METRIC = ("unknowns solved/s + % HBM roofline; served p50/p99 ms vs plain, "
          "same accuracy")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--workload", default="uniform",
                   choices=["uniform", "ragged", "serve", "gauss-sweep", "pivoting", "adi"],
                   help="uniform: 65,536 x 1,024 by default; ragged: 100,000 systems of "
                        "log-uniform sizes in [16, 8192], FP32 solve + FP64 refinement; "
                        "serve: request server with Poisson arrivals and resident ADI grids; "
                        "pivoting: the partial-pivoting (LAPACK ?gtsv) solve of a batch "
                        "that needs row interchanges; adi: Peaceman-Rachford steps of a "
                        "resident 2048^2 and 4096^2 FP64 grid")
    p.add_argument("--requests", type=int, default=64)
    p.add_argument("--served-requests", type=int, default=24,
                   help="uniform workload: requests of the served leg added to the line "
                        "(the request server with Poisson arrivals and resident ADI "
                        "grids); 0 = no served leg")
    p.add_argument("--ragged-mode", default="tol", choices=["tol", "refine", "fp64"],
                   help="ragged: FP32 solve + FP64 refinements until every system's "
                        "relative residual is <= 1e-12 (one or two corrections; the mixed-precision "
                        "workload), FP32 solve + exactly 2 FP64 refinements, or one FP64 "
                        "solve per system (all are timed; this picks the reported value)")
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--method", default="crpcr")
    p.add_argument("--n", type=int, default=1024)
    p.add_argument("--systems", type=int, default=65536)
    p.add_argument("--rest", type=int, default=0,
                   help="crpcr: CR->PCR switch size (0 = the plan's choice, one PCR "
                        "equation per lane: 64 at 1,024 unknowns); pcrgauss: G")
    p.add_argument("--dtype", default="f32", choices=["f32", "f64"])
    p.add_argument("--batch-stride", type=int, default=0,
                   help="uniform batches: system s at s * stride (>= n) through "
                        "trid_solve_strided (rocSPARSE's strided-batch layout); timing only")
    p.add_argument("--kind", default="dominant", choices=["dominant", "weak"],
                   help="uniform batches: strongly (b = |a|+|c|+U[1,2]) or weakly "
                        "(b = |a|+|c|+U[0.01,0.1]) diagonally dominant inputs")
    p.add_argument("--lanes", type=int, default=0)
    p.add_argument("--per-lane", type=int, default=0)
    p.add_argument("--io", type=int, default=0)
    p.add_argument("--lds-pad", type=int, default=0,
                   help="LDS-resident methods: pad period (-1 = none, 0 = default)")
    p.add_argument("--two-pass", action="store_true", help="pcrgauss: force the two-pass form")
    p.add_argument("--depths", default="2-10",
                   help="gauss-sweep: PCR depths to run, a range 'a-b' or a list 'a,b,...' "
                        "(2..8 form the sweep, 9 and 10 are reported beside it)")
    p.add_argument("--cpu-seconds", type=float, default=10.0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--layout", default="system", choices=["system", "interleaved"],
                   help="uniform workload: system-major arrays (the reference layout) or "
                        "interleaved (row i of system j at i*N + j, rocSPARSE's "
                        "gtsv_interleaved_batch layout)")
    p.add_argument("--no-vendor", action="store_true",
                   help="skip the rocSPARSE comparator and the device-copy ceiling")
    p.add_argument("--copy-probe", action="store_true",
                   help="after timing, copy the four input arrays once with a plain "
                        "device copy (known byte count) to calibrate HBM counters")
    return p.parse_args()


def traffic_from_profiles(workload_key: str):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by a copy kernel of known size), if committed."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(workload_key, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def vendor_baseline(a, b, c, d, x_ours, n, N, dtype, stream, reps=20):
    """rocSPARSE's gtsv_no_pivot_strided_batch (the vendor library's batched
    no-pivot solver over the same system-major layout) on the same inputs,
    timed with HIP events around the library call only (it overwrites its
    right-hand side, so d is copied into the output before every call,
    outside the timed interval); also the largest relative difference from
    this library's result."""
    import ctypes

    import torch

    try:
        rs = ctypes.CDLL("librocsparse.so")
    except OSError as e:
        return {"error": f"librocsparse.so not loadable: {e}"}
    pre = "s" if dtype == "f32" else "d"
    vp = ctypes.c_void_p
    hdl = vp()
    if rs.rocsparse_create_handle(ctypes.byref(hdl)) != 0:
        return {"error": "rocsparse_create_handle failed"}
    rs.rocsparse_set_stream(hdl, vp(stream.cuda_stream))
    xv = torch.empty_like(d)
    args_ = (hdl, ctypes.c_int(n), vp(a.data_ptr()), vp(b.data_ptr()), vp(c.data_ptr()),
             vp(xv.data_ptr()), ctypes.c_int(N), ctypes.c_int(n))
    size = ctypes.c_size_t(0)
    st = getattr(rs, f"rocsparse_{pre}gtsv_no_pivot_strided_batch_buffer_size")(
        *args_, ctypes.byref(size))
    if st != 0:
        rs.rocsparse_destroy_handle(hdl)
        return {"error": f"buffer_size status {st}"}
    buf = torch.empty(max(1, size.value), dtype=torch.uint8, device=d.device)
    solve = getattr(rs, f"rocsparse_{pre}gtsv_no_pivot_strided_batch")
    times = []
    for i in range(reps + 2):
        xv.copy_(d)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        st = solve(*args_, vp(buf.data_ptr()))
        e1.record(stream)
        if st != 0:
            rs.rocsparse_destroy_handle(hdl)
            return {"error": f"solve status {st}"}
        torch.cuda.synchronize()
        if i >= 2:
            times.append(e0.elapsed_time(e1))
    rs.rocsparse_destroy_handle(hdl)
    ms = sum(times) / len(times)
    diff = ((xv - x_ours).abs().view(N, n).amax(dim=1) /
            x_ours.abs().view(N, n).amax(dim=1)).max().item()
    esize = 4 if dtype == "f32" else 8
    return {"library": f"rocSPARSE rocsparse_{pre}gtsv_no_pivot_strided_batch",
            "kernel_ms": ms, "unknowns_per_s": n * N / (ms * 1e-3),
            "frac_of_peak": 5 * n * N * esize / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
            "max_rel_diff_from_ours": diff}


def bench_interleaved(args):
    """Uniform batch in the interleaved layout: the seeded system-major batch
    transposed on the device, solved by trid_solve_interleaved, and (unless
    --no-vendor) by rocSPARSE's gtsv_interleaved_batch (Thomas and the
    default QR algorithm) on copies of the same arrays."""
    import ctypes

    import torch

    import tridiagonalsolvers_amd as T

    n, N = args.n, args.systems
    tdt = torch.float32 if args.dtype == "f32" else torch.float64
    esize = 4 if args.dtype == "f32" else 8
    stream = torch.cuda.current_stream()
    sm = [torch.empty(n * N, dtype=tdt, device="cuda") for _ in range(4)]
    T.generate(0, 1, *sm, n, N, stream=stream)
    a, b, c, d = (v.view(N, n).t().contiguous().view(-1) for v in sm)
    x = torch.empty_like(a)
    h = T.Handle()

    def step():
        h.solve_interleaved(a, b, c, d, x, n, N, stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    t0 = time.perf_counter()
    for e0, e1 in evs:
        e0.record(stream)
        step()
        e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms = [e0.elapsed_time(e1) for e0, e1 in evs]
    avg = sum(kms) / len(kms)
    alg = 5 * n * N * esize
    # smallest system size (exclusive) of the library's register form for
    # interleaved batches: 256 for FP32, 128 for FP64
    il_lo = 256 if args.dtype == "f32" else 128
    out = {"metric": METRIC, "value": n * N * args.steps / wall, "unit": "unknowns/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": args.dtype,
           "data": "synthetic (seeded device generator, transposed to the interleaved layout)",
           "config": {"workload": f"interleaved {args.dtype} {N} systems x {n} unknowns",
                      "systems": N, "system_size": n},
           "kernel_ms": {"avg": avg, "min": min(kms)},
           "roofline": {"bound": "hbm", "achieved": alg / (avg * 1e-3) / 1e9,
                        "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": alg / (avg * 1e-3) / 1e9 / PEAK_HBM_GBS,
                        "traffic": traffic_from_profiles(
                            f"interleaved {args.dtype} {N} systems x {n} unknowns"),
                        "note": (f"{il_lo} < n <= 1024: register CR-PCR on systems gathered "
                                 "through LDS, 5 x T bytes per unknown moved"
                                 if il_lo < n <= 1024 else
                                 "the two-ended elimination moves 9 x T bytes per unknown "
                                 "(reads a, b, c, d; writes and re-reads c', d'; writes x)")}}
    if not args.no_vendor:
        rs = ctypes.CDLL("librocsparse.so")
        vp = ctypes.c_void_p
        pre = "s" if args.dtype == "f32" else "d"
        hdl = vp()
        rs.rocsparse_create_handle(ctypes.byref(hdl))
        rs.rocsparse_set_stream(hdl, vp(stream.cuda_stream))
        vend = {}
        for alg_id, name in ((1, "thomas"), (0, "default (QR)")):
            cp = [v.clone() for v in (a, b, c, d)]
            size = ctypes.c_size_t(0)
            st = getattr(rs, f"rocsparse_{pre}gtsv_interleaved_batch_buffer_size")(
                hdl, ctypes.c_int(alg_id), ctypes.c_int(n), *(vp(v.data_ptr()) for v in cp),
                ctypes.c_int(N), ctypes.c_int(N), ctypes.byref(size))
            if st != 0:
                vend[name] = {"error": f"buffer_size status {st}"}
                continue
            buf = torch.empty(max(1, size.value), dtype=torch.uint8, device="cuda")
            times = []
            for i in range(7):
                for dst, src in zip(cp, (a, b, c, d)):
                    dst.copy_(src)
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                st = getattr(rs, f"rocsparse_{pre}gtsv_interleaved_batch")(
                    hdl, ctypes.c_int(alg_id), ctypes.c_int(n), *(vp(v.data_ptr()) for v in cp),
                    ctypes.c_int(N), ctypes.c_int(N), vp(buf.data_ptr()))
                e1.record(stream)
                torch.cuda.synchronize()
                if st != 0:
                    break
                if i >= 2:
                    times.append(e0.elapsed_time(e1))
            if st != 0:
                vend[name] = {"error": f"solve status {st}"}
                continue
            ms = sum(times) / len(times)
            xv = cp[3]
            diff = ((xv - x).abs().view(n, N).amax(dim=0) /
                    x.abs().view(n, N).amax(dim=0)).max().item()
            vend[name] = {"kernel_ms": ms, "unknowns_per_s": n * N / (ms * 1e-3),
                          "max_rel_diff_from_ours": diff}
        rs.rocsparse_destroy_handle(hdl)
        out["vendor_baseline"] = {"library": f"rocSPARSE rocsparse_{pre}gtsv_interleaved_batch",
                                  **vend}
    if not args.no_cpu:
        xs = x.view(n, N).t().contiguous().view(-1)
        out["cpu_baseline"] = cpu_baseline(args, xs.cpu().numpy(), n, N)
    h.close()
    print(json.dumps(out), flush=True)


def bench_pivoting(args):
    """Partial-pivoting solve (trid_solve_pivoting_*, LAPACK ?gtsv
    semantics) of the seeded "paired" batch, which is not diagonally dominant
    and defeats the no-pivoting methods. System-major by default, or the
    interleaved layout with --layout interleaved. Unless --no-vendor, the same
    systems are also solved by rocSPARSE's gtsv_interleaved_batch with its
    LU algorithm (partial pivoting) in the interleaved layout, the vendor's
    batched pivoting solver. The CPU baseline is the oracle's ?gtsv
    restatement (OpenMP over all host cores); the GPU result is compared with
    it bit for bit."""
    import ctypes

    import numpy as np
    import torch

    import tridiagonalsolvers_amd as T

    n, N = args.n, args.systems
    tdt = torch.float32 if args.dtype == "f32" else torch.float64
    esize = 4 if args.dtype == "f32" else 8
    stream = torch.cuda.current_stream()
    sm = [torch.empty(n * N, dtype=tdt, device="cuda") for _ in range(4)]
    T.generate(3, 1, *sm, n, N, stream=stream)
    inter = args.layout == "interleaved"
    if inter:
        a, b, c, d = (v.view(N, n).t().contiguous().view(-1) for v in sm)
    else:
        a, b, c, d = sm
    x = torch.empty_like(a)
    info = torch.empty(N, dtype=torch.int32, device="cuda")
    h = T.Handle()

    def step():
        h.solve_pivoting(a, b, c, d, x, n, N, info=info,
                         batchStride=N if inter else None, stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    t0 = time.perf_counter()
    for e0, e1 in evs:
        e0.record(stream)
        step()
        e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms = [e0.elapsed_time(e1) for e0, e1 in evs]
    avg = sum(kms) / len(kms)
    alg = 5 * n * N * esize
    layout = "interleaved" if inter else "system-major"
    workload = f"pivoting {layout} {args.dtype} {N} systems x {n} unknowns"
    out = {"metric": METRIC, "value": n * N * args.steps / wall, "unit": "unknowns/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": args.dtype,
           "data": "synthetic (seeded device generator, kind 3: row pairs coupled by "
                   "1+U, diagonals 0.001 U[-1,1]; needs row interchanges)",
           "config": {"workload": workload, "systems": N, "system_size": n,
                      "layout": layout},
           "kernel_ms": {"avg": avg, "min": min(kms)},
           "singular_systems": int((info != 0).sum().item()),
           "roofline": {"bound": "hbm", "achieved": alg / (avg * 1e-3) / 1e9,
                        "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": alg / (avg * 1e-3) / 1e9 / PEAK_HBM_GBS,
                        "traffic": traffic_from_profiles(workload),
                        "note": "algorithmic bytes 5 x T per unknown; the checkpointed "
                                "two-pass elimination moves 9 x T (a, b, c, d read twice, "
                                "x written) plus 4 x T per chunk checkpoint"}}
    if not args.no_vendor:
        ia, ib, ic, idd = (v if inter else v.view(N, n).t().contiguous().view(-1)
                           for v in (a, b, c, d))
        xi = x if inter else x.view(N, n).t().contiguous().view(-1)
        rs = ctypes.CDLL("librocsparse.so")
        vp = ctypes.c_void_p
        pre = "s" if args.dtype == "f32" else "d"
        hdl = vp()
        rs.rocsparse_create_handle(ctypes.byref(hdl))
        rs.rocsparse_set_stream(hdl, vp(stream.cuda_stream))
        lu = 2  # rocsparse_gtsv_interleaved_alg_lu: LU with partial pivoting
        cp = [v.clone() for v in (ia, ib, ic, idd)]
        size = ctypes.c_size_t(0)
        st = getattr(rs, f"rocsparse_{pre}gtsv_interleaved_batch_buffer_size")(
            hdl, ctypes.c_int(lu), ctypes.c_int(n), *(vp(v.data_ptr()) for v in cp),
            ctypes.c_int(N), ctypes.c_int(N), ctypes.byref(size))
        vend = {"library": f"rocSPARSE rocsparse_{pre}gtsv_interleaved_batch (LU, "
                           f"partial pivoting, interleaved layout)"}
        if st != 0:
            vend["error"] = f"buffer_size status {st}"
        else:
            buf = torch.empty(max(1, size.value), dtype=torch.uint8, device="cuda")
            times = []
            for i in range(7):
                for dst, src in zip(cp, (ia, ib, ic, idd)):
                    dst.copy_(src)
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                st = getattr(rs, f"rocsparse_{pre}gtsv_interleaved_batch")(
                    hdl, ctypes.c_int(lu), ctypes.c_int(n), *(vp(v.data_ptr()) for v in cp),
                    ctypes.c_int(N), ctypes.c_int(N), vp(buf.data_ptr()))
                e1.record(stream)
                torch.cuda.synchronize()
                if st != 0:
                    break
                if i >= 2:
                    times.append(e0.elapsed_time(e1))
            if st != 0:
                vend["error"] = f"solve status {st}"
            else:
                ms = sum(times) / len(times)
                diff = ((cp[3] - xi).abs().view(n, N).amax(dim=0) /
                        xi.abs().view(n, N).amax(dim=0)).max().item()
                vend.update({"kernel_ms": ms, "unknowns_per_s": n * N / (ms * 1e-3),
                             "max_rel_diff_from_ours": diff})
        rs.rocsparse_destroy_handle(hdl)
        out["vendor_baseline"] = vend
    if not args.no_cpu:
        from oracle import oracle as O

        dt = np.float32 if args.dtype == "f32" else np.float64
        ha, hb, hc, hd = O.generate("paired", 1, n, N, dt)
        threads = O.max_threads()
        reps, t1 = 0, time.perf_counter()
        while True:
            xr, infor = O.gtsv(ha, hb, hc, hd, n, threads)
            reps += 1
            el = time.perf_counter() - t1
            if el >= args.cpu_seconds:
                break
        xs = (x.view(n, N).t().contiguous().view(-1) if inter else x).cpu().numpy()
        same = bool(np.all((xs == xr) | (np.isnan(xs) & np.isnan(xr))))
        out["cpu_baseline"] = {
            "value": reps * n * N / el, "unit": "unknowns/s", "cores": threads,
            "kind": "port",
            "sample": (f"oracle ?gtsv restatement ({args.dtype}), full {N}x{n} batch x "
                       f"{reps} passes, {el:.1f} s, OpenMP {threads} threads"),
            "gpu_bitwise_equal_to_oracle": same,
            "gpu_info_equal_to_oracle": bool(np.array_equal(info.cpu().numpy(), infor)),
        }
    h.close()
    print(json.dumps(out), flush=True)


def bench_adi(args):
    """N4: Peaceman-Rachford ADI steps (r = 1.5) of the 2D heat equation on
    resident FP64 grids of 2048^2 and 4096^2, all timed steps in one
    trid_adi_step call (per step: row rhs, row solve, column rhs transposed,
    column solve -- 4 passes of 16 bytes per point, u' staying transposed
    between steps; one transpose back per call). The 4096^2 grid is the
    reported workload. Unless
    --no-cpu, one step on a random grid of each size is checked against the
    oracle's FP64 ADI step (max relative error <= 1e-12, else the run fails)
    and the oracle step on 2048^2 is timed on the host as the CPU baseline."""
    import numpy as np
    import torch

    from tridiagonalsolvers_amd.adi import AdiGrid

    stream = torch.cuda.current_stream()
    per = {}
    for n in (2048, 4096):
        u = torch.rand(n * n, dtype=torch.float64, device="cuda",
                       generator=torch.Generator(device="cuda").manual_seed(n))
        g = AdiGrid(n, u)
        g.step(1.5, max(1, args.warmup), stream=stream)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        g.step(1.5, args.steps, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / args.steps
        ms = e0.elapsed_time(e1) / args.steps
        g.close()
        per[n] = {"step_ms": ms, "wall_ms_per_step": wall * 1e3,
                  "points_per_s": n * n / (ms * 1e-3),
                  "hbm_gbs_64B_per_point": 64 * n * n / (ms * 1e-3) / 1e9}
    n = 4096
    out = {"metric": METRIC, "value": n * n / (per[n]["wall_ms_per_step"] * 1e-3),
           "unit": "grid points/s (one ADI step = two sweeps of n systems of n)",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": per[n]["wall_ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (U[0,1) grids, seeded torch generator), r = 1.5",
           "config": {"workload": "adi f64 4096^2 resident grid, Peaceman-Rachford step",
                      "grids": {str(k): v for k, v in per.items()}},
           "roofline": {"bound": "hbm", "achieved": per[n]["hbm_gbs_64B_per_point"],
                        "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": per[n]["hbm_gbs_64B_per_point"] / PEAK_HBM_GBS,
                        "traffic": traffic_from_profiles(
                            "adi f64 4096^2 resident grid, Peaceman-Rachford step"),
                        "note": "64 B per point and step: 4 passes (row rhs, row solve, "
                                "column rhs + transpose, column solve) of 16 B each; one "
                                "transpose back per call is not credited"}}
    if not args.no_cpu:
        from oracle import oracle as O

        # one step on each grid size against the oracle's FP64 step
        errs = {}
        for m in (2048, 4096):
            u0 = np.random.default_rng(m + 5).random(m * m)
            dev = torch.from_numpy(u0).cuda()
            g = AdiGrid(m, dev)
            g.step(1.5, 1, stream=stream)
            torch.cuda.synchronize()
            ref = O.adi_step_np(u0.reshape(m, m), 1.5)
            got = dev.cpu().numpy().reshape(m, m)
            errs[m] = float(np.abs(got - ref).max() / np.abs(ref).max())
            g.close()
        m = 2048
        u0 = np.random.default_rng(5).random(m * m)
        reps, t1 = 0, time.perf_counter()
        while True:
            ref = O.adi_step_np(u0.reshape(m, m), 1.5)
            reps += 1
            el = time.perf_counter() - t1
            if el >= args.cpu_seconds:
                break
        out["cpu_baseline"] = {
            "value": reps * m * m / el, "unit": "grid points/s", "cores": O.max_threads(),
            "kind": "port",
            "sample": f"oracle FP64 ADI step on 2048^2 x {reps}, {el:.1f} s (Thomas per "
                      f"sweep, OpenMP over rows)",
            "gpu_max_rel_err_vs_oracle_step": max(errs.values()),
            "gpu_max_rel_err_vs_oracle_step_by_grid": {str(k): v for k, v in errs.items()}}
        if max(errs.values()) > 1e-12:
            raise SystemExit(f"ADI step differs from the oracle step: {errs}")
    print(json.dumps(out), flush=True)


def copy_ceiling(a, b, c, d, stream, reps=10):
    """Bandwidth of a plain device copy of the four input arrays (torch's copy
    kernel): read + write bytes over HIP-event time, the achievable HBM rate
    to compare roofline fractions against."""
    import torch

    dst = torch.empty_like(a)
    for _ in range(2):
        dst.copy_(a)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        for src in (a, b, c, d):
            dst.copy_(src)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    moved = 2 * 4 * a.numel() * a.element_size()
    return {"gbs": moved / (ms * 1e-3) / 1e9, "frac_of_peak": moved / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
            "bytes": moved, "ms": ms}


def stream_ceiling(a, b, c, d, x, stream, reps=20):
    """Time of a plain stream of the solve's bytes (read a, b, c, d; write x;
    16-byte non-temporal accesses, no arithmetic; bench_stream.so, built by
    __graft_entry__.build()): the speed of light of this traffic on the
    device, against which the solve kernel's time is read."""
    import ctypes

    import torch

    path = os.path.join(ROOT, "bench_stream.so")
    if not os.path.exists(path):
        return {"error": "bench_stream.so not built"}
    lib = ctypes.CDLL(path)
    f = lib.bench_stream_4r1w
    f.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_int64, ctypes.c_void_p]
    f.restype = ctypes.c_int
    nbytes = a.numel() * a.element_size()
    if nbytes % 16 or any(v.data_ptr() % 16 for v in (a, b, c, d, x)):
        return {"error": "arrays not 16-byte aligned"}
    args = [ctypes.c_void_p(v.data_ptr()) for v in (a, b, c, d)]
    dst = torch.empty_like(x)
    call = lambda: f(*args, ctypes.c_void_p(dst.data_ptr()), nbytes // 16,
                     ctypes.c_void_p(stream.cuda_stream))
    for _ in range(3):
        if call() != 0:
            return {"error": "launch failed"}
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        call()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    moved = 5 * nbytes
    return {"us": ms * 1e3, "gbs": moved / (ms * 1e-3) / 1e9,
            "frac_of_peak": moved / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, "bytes": moved}


def cpu_baseline(args, x_gpu_np, n, N):
    """The oracle's Thomas (same recurrence as tridiagonal_gold.h:26-44) with
    OpenMP over all host cores on the same inputs: the FP64 recurrence (the
    oracle itself, on the FP32-rounded inputs of an FP32 run) is the reported
    value, and for FP32 runs the FP32 recurrence is timed beside it for a
    like-for-like figure; each is repeated for about half of --cpu-seconds.
    Also reports the GPU result's max relative error against the FP64
    oracle on those inputs."""
    import numpy as np

    from oracle import oracle as O

    dt = np.float32 if args.dtype == "f32" else np.float64
    a, b, c, d = O.generate(getattr(args, "kind", "dominant"), 1, n, N, dt)
    threads = O.max_threads()

    def timed(fn, seconds):
        reps, t0 = 0, time.perf_counter()
        while True:
            fn()
            reps += 1
            el = time.perf_counter() - t0
            if el >= seconds:
                return reps, el

    share = args.cpu_seconds / (2.0 if dt == np.float32 else 1.0)
    reps64, el64 = timed(lambda: O.thomas(a, b, c, d, n=n, threads=threads), share)
    xref = O.thomas(a, b, c, d, n=n, threads=threads)
    rel, rms, fails = O.compare(x_gpu_np, xref, n=n, rms_eps=1e-3)
    out = {
        "value": reps64 * n * N / el64,
        "unit": "unknowns/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle Thomas in FP64 arithmetic ({args.dtype} inputs), full {N}x{n} "
                   f"batch x {reps64} passes, {el64:.1f} s, OpenMP {threads} threads"),
        "gpu_max_rel_err_vs_fp64_thomas": float(np.max(rel)),
        "gpu_reference_rms_failures": int(fails),
    }
    if dt == np.float32:
        reps32, el32 = timed(lambda: O.thomas_f32(a, b, c, d, n, threads), share)
        out["t_precision"] = {
            "value": reps32 * n * N / el32, "unit": "unknowns/s", "cores": threads,
            "sample": (f"oracle Thomas in FP32 arithmetic, full {N}x{n} batch x {reps32} "
                       f"passes, {el32:.1f} s, OpenMP {threads} threads")}
    return out


def ragged_cpu_baseline(args, host, offs, x_gpu):
    """FP64 Thomas of the oracle on the same ragged batch, OpenMP over all host
    cores, repeated for about --cpu-seconds."""
    import numpy as np

    from oracle import oracle as O

    a, b, c, d = host
    threads = O.max_threads()
    reps, t0 = 0, time.perf_counter()
    while True:
        xref = O.thomas(a, b, c, d, offsets=offs, threads=threads)
        reps += 1
        el = time.perf_counter() - t0
        if el >= args.cpu_seconds:
            break
    rel, _, fails = O.compare(x_gpu, xref, offsets=offs)
    res = O.residual(a, b, c, d, x_gpu, offsets=offs)
    return {"value": reps * int(offs[-1]) / el, "unit": "unknowns/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle FP64 Thomas, full ragged batch x {reps} passes, {el:.1f} s, "
                      f"OpenMP {threads} threads",
            "gpu_max_rel_err_vs_fp64_thomas": float(np.max(rel)),
            "gpu_max_rel_residual_host_long_double": float(np.max(res)),
            "gpu_reference_rms_failures": int(fails)}


def bench_ragged(args):
    """C4: 100,000 systems, sizes round(exp(U[ln 16, ln 8192])) (seed 1), FP64
    data; one graph replay of the persistent ragged kernel against plain FP64
    launches, one per distinct size, on the same data (sorted by size so each
    size is contiguous). The ragged kernel is timed in each of its modes: FP32
    solve + FP64 refinements to a relative residual of 1e-12 (at most 2),
    FP32 solve + exactly 2 or 1 FP64 refinements, one FP64 solve;
    --ragged-mode picks the reported value."""
    import numpy as np
    import torch

    import tridiagonalsolvers_amd as T
    import tridiagonalsolvers_amd.ragged as R

    sizes = np.sort(R.log_uniform_sizes(100000, 16, 8192, seed=1))
    offs = R.offsets_of(sizes)
    total = int(offs[-1])
    N = len(sizes)
    dev = torch.device("cuda:0")
    a, b, c, d = (torch.empty(total, dtype=torch.float64, device=dev) for _ in range(4))
    x = torch.empty_like(a)
    x_plain = torch.empty_like(a)
    stream = torch.cuda.current_stream()
    # every row strongly dominant: a,c ~ U[-1,1], b = |a|+|c|+U[1,2]; the
    # first/last rows' a/c are ignored by the solvers
    T.generate(0, 1, a, b, c, d, total, 1, stream=stream)
    plan_refine = R.RaggedPlan(offs)
    plan_fp64 = R.RaggedPlan(offs, table="default-fp64")
    info = (plan_fp64 if args.ragged_mode == "fp64" else plan_refine).info()
    h = T.Handle()
    uniq, first = np.unique(sizes, return_index=True)
    runs = list(zip(uniq.tolist(), first.tolist(),
                    np.diff(np.append(first, N)).tolist()))
    TOL, TOL_ITERS = 1e-12, 2
    names = {"tol": f"refine to 1e-12 (FP32 solve + FP64 refinements until every system's "
                    f"relative residual is <= {TOL:g}, at most {TOL_ITERS})",
             "refine": "refine (FP32 solve + 2 FP64 refinements)",
             "refine1": "refine (FP32 solve + 1 FP64 refinement)",
             "fp64": "fp64 (one FP64 solve)"}
    steps_of = {
        "tol": lambda xo: plan_refine.solve(a, b, c, d, xo, iters=TOL_ITERS, tol=TOL,
                                            stream=stream),
        "refine": lambda xo: plan_refine.solve(a, b, c, d, xo, iters=2, stream=stream),
        "refine1": lambda xo: plan_refine.solve(a, b, c, d, xo, iters=1, stream=stream),
        "fp64": lambda xo: plan_fp64.solve(a, b, c, d, xo, stream=stream, fp64=True)}

    def plain_step():
        for n, f, cnt in runs:
            o = int(offs[f]) * 8
            h.solve("crpcr", a.data_ptr() + o, b.data_ptr() + o, c.data_ptr() + o,
                    d.data_ptr() + o, x_plain.data_ptr() + o, n, cnt, stream=stream,
                    dtype="f64")

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps, e0.elapsed_time(e1) / steps

    d_offs = torch.from_numpy(offs).to(dev)
    rr = torch.empty(N, dtype=torch.float64, device=dev)
    modes, timing = {}, {}
    main = args.ragged_mode
    for m in [main] + [k for k in steps_of if k != main]:
        xo = x if m == main else torch.empty_like(a)
        wall_m, gpu_m = timed(lambda: steps_of[m](xo), args.steps, args.warmup)
        R.residual(a, b, c, d, xo, d_offs, N, rr, stream=stream)
        timing[m] = (wall_m, gpu_m)
        modes[names[m]] = {"value": total / wall_m, "kernel_ms": gpu_m,
                           "max_rel_residual": float(rr.max())}
    wall_r, gpu_r = timing[main]
    rr_r = modes[names[main]]["max_rel_residual"]
    fp64 = main == "fp64"
    wall_p, gpu_p = timed(plain_step, max(1, args.steps // 5), 1)
    R.residual(a, b, c, d, x_plain, d_offs, N, rr, stream=stream)
    rr_p = float(rr.max())
    alg = 5 * 8 * total
    out = {
        "metric": METRIC, "value": total / wall_r, "unit": "unknowns/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall_r * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded device generator, FP64; sizes log-uniform in [16, 8192], seed 1)",
        "config": {"workload": f"ragged f64 {N} systems, sizes log-uniform [16, 8192], "
                               f"{names[main]}, one graph launch",
                   "systems": N, "unknowns": total, "plan": info},
        "modes": modes,
        "kernel_ms": {"ragged_graph": gpu_r, "plain_per_size": gpu_p},
        "roofline": {"bound": "hbm", "achieved": alg / (gpu_r * 1e-3) / 1e9,
                     "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": alg / (gpu_r * 1e-3) / 1e9 / PEAK_HBM_GBS,
                     "traffic": traffic_from_profiles(
                         f"ragged f64 {N} systems, sizes log-uniform [16, 8192], "
                         f"{names[main]}, one graph launch"),
                     "note": "algorithmic 5*8 B per unknown (FP64 in/out)"
                             + ("" if fp64 else "; refinement re-reads are not credited")},
        "plain_per_size": {"value": total / wall_p, "launches": len(runs),
                           "ms_per_step": wall_p * 1e3, "max_rel_residual": rr_p},
        "speedup_vs_plain_per_size": wall_p / wall_r,
        "max_rel_residual": rr_r,
    }
    if not args.no_cpu:
        host = [v.cpu().numpy() for v in (a, b, c, d)]
        out["cpu_baseline"] = ragged_cpu_baseline(args, host, offs, x.cpu().numpy())
    plan_refine.close()
    plan_fp64.close()
    h.close()
    print(json.dumps(out), flush=True)


def _percentile(v, q):
    import numpy as np
    return float(np.percentile(np.asarray(v, dtype=np.float64), q)) if len(v) else None


def serve_measure(n_requests: int, pageable: bool = True, poisson_frac: float = 0.7):
    """C5 on the request server: `n_requests` requests of 1..4,096 systems
    with sizes log-uniform in [32, 16,384] (FP64 host data, FP32 solve + 2
    FP64 refinements), slices of one seeded pool. Measures the served rate at
    saturation (all requests submitted at once, arrays in the server's
    page-locked memory, and with `pageable` also from ordinary memory), the
    same requests as plain per-size FP64 launches back to back, and p50/p99
    latency of Poisson arrivals at `poisson_frac` of the saturated request
    rate, interleaved with ADI steps (r ~ U[0.1, 10]) on resident 2048^2 and
    4096^2 grids. The relative residual of every served request is computed
    on the device. Returns a dict of the figures."""
    import threading

    import numpy as np
    import torch

    import tridiagonalsolvers_amd as T
    import tridiagonalsolvers_amd.ragged as R
    import tridiagonalsolvers_amd.server as S

    rng = np.random.default_rng(1)
    max_req = 1 << 24
    pool_n = 1 << 25
    dev = torch.device("cuda:0")
    pool = [torch.empty(pool_n, dtype=torch.float64, device=dev) for _ in range(4)]
    T.generate(0, 7, *pool, pool_n, 1)
    host_pageable = [v.cpu().numpy() for v in pool]
    host = host_pageable
    reqs = []
    while len(reqs) < n_requests:
        k = int(rng.integers(1, 4097))
        sizes = R.log_uniform_sizes(k, 32, 16384, seed=int(rng.integers(1 << 30)))
        tot = int(sizes.sum())
        if tot > max_req:
            continue
        start = int(rng.integers(0, pool_n - tot))
        reqs.append((R.offsets_of(sizes), start, tot))
    total_unknowns = sum(r[2] for r in reqs)

    def req_arrays(i):
        offs, start, tot = reqs[i]
        return offs, [h[start:start + tot] for h in host]

    srv = S.Server(hbm_budget_bytes=16 << 30, slots=3, max_batch_unknowns=max_req,
                   batch_window_us=300)
    xs = [np.empty(r[2]) for r in reqs]
    # warm-up (kernel load, first-touch of staging)
    offs0, arr0 = req_arrays(0)
    srv.wait(srv.submit(arr0[3], xs[0], offs0, *arr0[:3]))

    # --- saturation: everything submitted at once -----------------------------
    def saturate():
        t0 = time.perf_counter()
        tickets = []
        for i in range(len(reqs)):
            offs, arr = req_arrays(i)
            tickets.append(srv.submit(arr[3], xs[i], offs, *arr[:3]))
        sat = [srv.wait(t) for t in tickets]
        el = time.perf_counter() - t0
        assert all(r["state"] == "done" for r in sat)
        return el

    # client arrays in ordinary (pageable) memory: staged through pinned
    # buffers by the server
    sat_page_s = saturate() if pageable else None
    # client arrays in the server's page-locked memory: DMA with no staging
    host = [srv.host_array(pool_n) for _ in range(4)]
    for dst, src in zip(host, host_pageable):
        dst[:] = src
    xs_flat = srv.host_array(total_unknowns)
    xs, pos = [], 0
    for r in reqs:
        xs.append(xs_flat[pos:pos + r[2]])
        pos += r[2]
    saturate()            # first touch of the new buffers
    sat_s = saturate()
    served_rate = total_unknowns / sat_s

    # accuracy of the served results: device residual of every request
    worst_res = 0.0
    for i in range(len(reqs)):
        offs, start, tot = reqs[i]
        d_offs = torch.from_numpy(offs).to(dev)
        x = torch.from_numpy(xs[i]).to(dev)
        rr = torch.empty(len(offs) - 1, dtype=torch.float64, device=dev)
        R.residual(*(v[start:start + tot] for v in pool), x, d_offs, len(offs) - 1, rr)
        worst_res = max(worst_res, float(rr.max()))
    del pool

    # --- plain per-size FP64 launches of the same requests, back to back ----------
    h = T.Handle()
    x_dev = torch.empty(max_req, dtype=torch.float64, device=dev)
    t0 = time.perf_counter()
    for i in range(len(reqs)):
        offs, arr = req_arrays(i)
        sizes = np.diff(offs)
        order = np.argsort(sizes, kind="stable")
        idx = np.concatenate([np.arange(offs[s], offs[s + 1]) for s in order])
        g = [torch.from_numpy(v[idx]).to(dev) for v in arr]
        ssz = sizes[order]
        uniq, first, cnt = np.unique(ssz, return_index=True, return_counts=True)
        pos = np.concatenate([[0], np.cumsum(ssz)])
        for n, f, c in zip(uniq.tolist(), first.tolist(), cnt.tolist()):
            o = int(pos[f]) * 8
            # FP64 CR-PCR holds systems up to 8192 in one workgroup; longer
            # ones use the two-pass PCR + Thomas form
            method = "crpcr" if n <= 8192 else "pcrgauss"
            h.solve(method, *(v.data_ptr() + o for v in g), x_dev.data_ptr() + o, n, c,
                    dtype="f64")
        out = np.empty(len(idx))
        out[idx] = x_dev[:len(idx)].cpu().numpy()
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t0
    plain_rate = total_unknowns / plain_s
    h.close()

    # --- Poisson arrivals at a fraction of the saturated request rate, with ADI -
    lam = poisson_frac * len(reqs) / sat_s
    gaps = rng.exponential(1.0 / lam, size=len(reqs))
    grids = {}
    for n in (2048, 4096):
        grids[n] = srv.create_adi(np.zeros(n * n) + 1.0)
    adi_period = 4.0 / lam
    adi_lat = {2048: [], 4096: []}
    stop = threading.Event()

    def adi_client():
        r_rng = np.random.default_rng(2)
        tk = []
        while not stop.is_set():
            for n, gid in grids.items():
                tk.append((n, srv.submit_adi(gid, float(r_rng.uniform(0.1, 10.0)))))
            time.sleep(adi_period)
        for n, t in tk:
            adi_lat[n].append(srv.wait(t)["latency_us"])

    th = threading.Thread(target=adi_client)
    th.start()
    tickets = []
    t0 = time.perf_counter()
    for i, gap in enumerate(gaps):
        time.sleep(gap)
        offs, arr = req_arrays(i)
        tickets.append(srv.submit(arr[3], xs[i], offs, *arr[:3]))
    res = [srv.wait(t) for t in tickets]
    poisson_s = time.perf_counter() - t0
    stop.set()
    th.join()
    lat = [r["latency_us"] / 1e3 for r in res]
    st = srv.stats()
    for gid in grids.values():
        srv.release_adi(gid)
    srv.close()

    def pct(v, q):
        return _percentile(v, q) if v else None

    return {
        "requests": len(reqs), "unknowns": total_unknowns, "slots": 3,
        "budget_bytes": 16 << 30,
        "sat_s": sat_s, "served_rate": served_rate, "sat_page_s": sat_page_s,
        "max_rel_residual": worst_res, "plain_s": plain_s, "plain_rate": plain_rate,
        "lam": lam, "poisson_s": poisson_s,
        "latency_ms_p50": pct(lat, 50), "latency_ms_p99": pct(lat, 99),
        "adi_2048_step_ms_p50": pct([v / 1e3 for v in adi_lat[2048]], 50),
        "adi_2048_step_ms_p99": pct([v / 1e3 for v in adi_lat[2048]], 99),
        "adi_4096_step_ms_p50": pct([v / 1e3 for v in adi_lat[4096]], 50),
        "adi_4096_step_ms_p99": pct([v / 1e3 for v in adi_lat[4096]], 99),
        "adi_steps": sum(len(v) for v in adi_lat.values()),
        "server_stats": st,
    }


SERVE_WORKLOAD = ("served: requests of 1..4096 systems, sizes log-uniform [32, 16384], "
                  "FP32 solve + 2 FP64 refinements; ADI 2048^2 and 4096^2 resident")


def served_summary(m: dict) -> dict:
    """The compact served object of the default benchmark line."""
    return {
        "workload": SERVE_WORKLOAD, "requests": m["requests"], "unknowns": m["unknowns"],
        "unknowns_per_s": m["served_rate"],
        "plain_per_size_unknowns_per_s": m["plain_rate"],
        "speedup_vs_plain_per_size": m["plain_s"] / m["sat_s"],
        "poisson_rate_req_per_s": m["lam"],
        "latency_ms_p50": m["latency_ms_p50"], "latency_ms_p99": m["latency_ms_p99"],
        "adi_2048_step_ms_p50": m["adi_2048_step_ms_p50"],
        "adi_2048_step_ms_p99": m["adi_2048_step_ms_p99"],
        "adi_4096_step_ms_p50": m["adi_4096_step_ms_p50"],
        "adi_4096_step_ms_p99": m["adi_4096_step_ms_p99"],
        "max_rel_residual": m["max_rel_residual"],
        "note": "host FP64 arrays in the server's page-locked memory, PCIe copies included; "
                "plain = the same requests as per-size FP64 CR-PCR launches (FP64 accuracy)",
    }


def bench_serve(args):
    """C5: the request server (serve_measure) with --requests requests; one
    JSON line, value = the served unknowns/s at saturation."""
    m = serve_measure(args.requests)
    out = {
        "metric": METRIC, "value": m["served_rate"], "unit": "unknowns/s", "n_gpus": 1,
        "steps": m["requests"], "warmup": 1, "ms_per_step": m["sat_s"] * 1e3 / m["requests"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded device generator, FP64 host copies; request sizes seeded)",
        "config": {"workload": f"{SERVE_WORKLOAD} ({m['requests']} requests)",
                   "unknowns": m["unknowns"], "slots": m["slots"],
                   "budget_bytes": m["budget_bytes"]},
        "served_saturated": {"unknowns_per_s": m["served_rate"], "seconds": m["sat_s"],
                             "max_rel_residual": m["max_rel_residual"],
                             "client_memory": "server page-locked buffers (DMA, no staging)"},
        "served_saturated_pageable": {"unknowns_per_s": m["unknowns"] / m["sat_page_s"],
                                      "seconds": m["sat_page_s"],
                                      "client_memory": "pageable (staged by the server)"},
        "plain_per_size": {"unknowns_per_s": m["plain_rate"], "seconds": m["plain_s"]},
        "speedup_vs_plain_per_size": m["plain_s"] / m["sat_s"],
        "poisson": {"rate_req_per_s": m["lam"], "seconds": m["poisson_s"],
                    "latency_ms_p50": m["latency_ms_p50"],
                    "latency_ms_p99": m["latency_ms_p99"],
                    "adi_2048_step_ms_p50": m["adi_2048_step_ms_p50"],
                    "adi_2048_step_ms_p99": m["adi_2048_step_ms_p99"],
                    "adi_4096_step_ms_p50": m["adi_4096_step_ms_p50"],
                    "adi_4096_step_ms_p99": m["adi_4096_step_ms_p99"],
                    "adi_steps": m["adi_steps"]},
        "server_stats": m["server_stats"],
    }
    print(json.dumps(out), flush=True)


def bench_gauss_sweep(args):
    """C3: 64 systems x 1,048,576 unknowns FP32, pcrGauss two-pass form
    (global PCR pass of depth k, then two-ended Gaussian elimination on the
    2^k interleaved subsystems), k = 2..8, plus k = 9, 10 for reference
    (not part of the sweep); unless --no-cpu every depth's result is
    compared with FP64 Thomas on the same inputs. One JSON line; value = the
    best depth's unknowns/s in 2..8. The same batch through crpcr (the
    separator form at this size) is reported beside it."""
    import numpy as np
    import torch

    import tridiagonalsolvers_amd as T

    n, N = 1 << 20, 64
    dev = torch.device("cuda:0")
    a, b, c, d = (torch.empty(n * N, dtype=torch.float32, device=dev) for _ in range(4))
    x = torch.empty_like(a)
    stream = torch.cuda.current_stream()
    T.generate(0, 1, a, b, c, d, n, N, stream=stream)
    h = T.Handle()
    xref = None
    if not args.no_cpu:
        # every depth's result against FP64 Thomas on the same inputs
        from oracle import oracle as O
        ah, bh, ch, dh = O.generate("dominant", 1, n, N, np.float32)
        xref = O.thomas(ah, bh, ch, dh, n=n, threads=O.max_threads())
        del ah, bh, ch, dh
    if "-" in args.depths:
        lo, hi = (int(v) for v in args.depths.split("-"))
        depths = list(range(lo, hi + 1))
    else:
        depths = [int(v) for v in args.depths.split(",")]
    if not depths or min(depths) < 1 or max(depths) > 19:
        raise SystemExit("--depths: PCR depths between 1 and 19")
    rows = []
    for k in depths:
        opts = T.make_opts(gaussian=1 << k)
        for _ in range(args.warmup):
            h.solve("pcrgauss", a, b, c, d, x, n, N, opts, stream=stream)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        steps = max(1, args.steps // 5)
        e0.record(stream)
        for _ in range(steps):
            h.solve("pcrgauss", a, b, c, d, x, n, N, opts, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        alg = 5 * 4 * n * N
        row = {"depth": k, "gaussian": 1 << k, "ms": ms, "unknowns_per_s": n * N / (ms * 1e-3),
               "alg_frac_of_peak": alg / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
               "moved_bytes_model": (7 + 8) * 4 * n * N,
               "in_sweep": k <= 8}
        if xref is not None:
            rel, _, fails = O.compare(x.cpu().numpy(), xref, n=n, rms_eps=1e-3)
            row["max_rel_err_vs_fp64_thomas"] = float(np.max(rel))
            row["reference_rms_failures"] = int(fails)
        rows.append(row)
    best = min((r for r in rows if r["in_sweep"]), key=lambda r: r["ms"], default=None)
    if best is None:
        best = min(rows, key=lambda r: r["ms"])
    # the same batch through crpcr, which takes the separator form at this
    # size (chunks of 1,024 rows; reads the inputs twice, writes x once)
    for _ in range(args.warmup):
        h.solve("crpcr", a, b, c, d, x, n, N, stream=stream)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        h.solve("crpcr", a, b, c, d, x, n, N, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    sep_ms = e0.elapsed_time(e1) / args.steps
    sep = {"method": "crpcr", "family": T.plan_query("crpcr", n, N, 4)["family"],
           "ms": sep_ms, "unknowns_per_s": n * N / (sep_ms * 1e-3),
           "alg_frac_of_peak": 5 * 4 * n * N / (sep_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
           "moved_bytes_model": 9 * 4 * n * N + (21 * 16 + 5 * 4) * (n // 1024) * N,
           "moved_frac_of_peak": (9 * 4 * n * N) / (sep_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
           "speedup_vs_best_depth": best["ms"] / sep_ms}
    if xref is not None:
        rel, _, fails = O.compare(x.cpu().numpy(), xref, n=n, rms_eps=1e-3)
        sep["max_rel_err_vs_fp64_thomas"] = float(np.max(rel))
        sep["reference_rms_failures"] = int(fails)
    out = {"metric": METRIC, "value": best["unknowns_per_s"], "unit": "unknowns/s", "n_gpus": 1,
           "steps": max(1, args.steps // 5), "warmup": args.warmup, "ms_per_step": best["ms"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded device generator)",
           "config": {"workload": "pcrgauss f32 64 systems x 1048576 unknowns, depth 2..8",
                      "best_depth": best["depth"]},
           "sweep": rows,
           "separator_form": sep,
           "roofline": {"bound": "hbm", "achieved": 5 * 4 * n * N / (best["ms"] * 1e-3) / 1e9,
                        "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": 5 * 4 * n * N / (best["ms"] * 1e-3) / 1e9 / PEAK_HBM_GBS,
                        "traffic": traffic_from_profiles(
                            f"pcrgauss f32 {N} systems x {n} unknowns, depth {best['depth']}"),
                        "note": "two passes move about 15 n x 4 B per system (pass 1 reads 4 "
                                "arrays and writes the 3 of the normalized rows, pass 2 reads "
                                "3, writes and re-reads the factors q', d', writes x), plus the "
                                "tiles' halo re-reads: at most 5/15 of peak in algorithmic "
                                "bytes"}}
    if not args.no_cpu:
        h.solve("pcrgauss", a, b, c, d, x, n, N, T.make_opts(gaussian=best["gaussian"]),
                stream=stream)
        torch.cuda.synchronize()
        out["cpu_baseline"] = cpu_baseline(args, x.cpu().numpy(), n, N)
    h.close()
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only; independent systems shard by system range "
              "with no exchange, so several GPUs each run this one-GPU workload.",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "ragged":
        bench_ragged(args)
        return
    if args.workload == "serve":
        bench_serve(args)
        return
    if args.workload == "gauss-sweep":
        bench_gauss_sweep(args)
        return
    if args.workload == "pivoting":
        bench_pivoting(args)
        return
    if args.workload == "adi":
        bench_adi(args)
        return
    if args.layout == "interleaved":
        bench_interleaved(args)
        return
    import torch

    import tridiagonalsolvers_amd as T

    n, N = args.n, args.systems
    tdt = torch.float32 if args.dtype == "f32" else torch.float64
    esize = 4 if args.dtype == "f32" else 8
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    bs = args.batch_stride if args.batch_stride > 0 else n
    if bs < n:
        raise SystemExit("--batch-stride must be at least --n")
    a, b, c, d = (torch.empty(bs * N, dtype=tdt, device=dev) for _ in range(4))
    x = torch.empty_like(a)
    stream = torch.cuda.current_stream()
    # strided: rows of systems of size bs, of which the first n form each system
    T.generate(0 if args.kind == "dominant" else 1, 1, a, b, c, d, bs, N, stream=stream)
    h = T.Handle()
    opts = T.make_opts(rest=args.rest if args.method in ("crpcr", "crpcrnbc") else 0,
                       gaussian=args.rest if args.method == "pcrgauss" else 0,
                       lanes=args.lanes, per_lane=args.per_lane, io_mode=args.io,
                       lds_pad=args.lds_pad, flags=1 if args.two_pass else 0)
    plan = T.plan_query(args.method, n, N, esize, opts)

    def step():
        if bs != n:
            h.solve_strided(args.method, a, b, c, d, x, n, N, bs, opts, stream=stream)
        else:
            h.solve(args.method, a, b, c, d, x, n, N, opts, stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # per-launch kernel time: HIP events on the launch stream
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1 in evs:
        e0.record(stream)
        step()
        e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms = sorted(e0.elapsed_time(e1) for e0, e1 in evs)
    avg_ms = sum(kms) / len(kms)
    med_ms = kms[len(kms) // 2]

    unknowns = n * N
    alg_bytes = 5 * n * esize * N
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    workload = f"{args.method} {args.dtype} {N} systems x {n} unknowns"
    if args.kind != "dominant":
        workload += f", {args.kind}ly dominant"
    if bs != n:
        workload += f", batch stride {bs}"
    traffic = traffic_from_profiles(workload)
    out = {
        "metric": METRIC,
        "value": unknowns * args.steps / wall,
        "unit": "unknowns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": ("synthetic (seeded device generator: a,c~U[-1,1], b=|a|+|c|+"
                 + ("U[1,2]" if args.kind == "dominant" else "U[0.01,0.1]") + ", d~U[-1,1])"),
        "config": {"workload": workload, "method": args.method, "systems": N,
                   "system_size": n, "rest": plan["rest"], "lanes": plan["lanes"],
                   "per_lane": plan["per_lane"], "family": plan["family"]},
        "kernel_ms": {"avg": avg_ms, "median": med_ms, "min": kms[0]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                     "traffic": traffic},
    }
    if bs != n:
        args.no_vendor = args.no_cpu = True  # the checks below assume a contiguous batch
    if not args.no_vendor:
        out["vendor_baseline"] = vendor_baseline(a, b, c, d, x, n, N, args.dtype, stream)
        out["copy_ceiling"] = copy_ceiling(a, b, c, d, stream)
        out["stream_ceiling"] = stream_ceiling(a, b, c, d, x, stream)
        if "us" in out["stream_ceiling"]:
            out["stream_ceiling"]["solve_time_over_stream_time"] = (
                avg_ms * 1e3 / out["stream_ceiling"]["us"])
    if args.copy_probe:
        # 4 * n * N * esize bytes read and written by a plain copy kernel
        dst = torch.empty_like(a)
        for src in (a, b, c, d):
            dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_probe_bytes_each_way"] = 4 * n * N * esize
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args, x.cpu().numpy(), n, N)
    h.close()
    if args.served_requests > 0 and bs == n:
        # the served half of the metric, on the same GPU in the same run
        del a, b, c, d, x
        torch.cuda.empty_cache()
        out["served"] = served_summary(serve_measure(args.served_requests, pageable=False))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
