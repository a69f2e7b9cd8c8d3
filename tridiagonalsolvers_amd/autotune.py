"""Per-bucket autotuner of the ragged solve.

For every padded size 2..16384 it times the candidate configurations of the
persistent ragged kernel -- unknowns per lane (threads per system), the
CR -> PCR switch size (`rest`, with rest = 2 being plain CR) -- on a batch of
that size alone, in the mode the table is for (FP32 solve + FP64
refinement, or one FP64 solve), and keeps the fastest. The result is a JSON
table (tune/ragged_f64_refine.json, tune/ragged_f64_fp64.json) that
RaggedPlan passes to trid_ragged_plan_create. Needs a gfx950 device.

    python -m tridiagonalsolvers_amd.autotune [--mode refine|fp64] [--out PATH]
"""
from __future__ import annotations

import argparse
import json
import os
import time


def candidates(size: int):
    out = []
    if size <= 512:
        m = 1
        while m <= 8:
            lanes = size // m
            if 2 <= lanes <= 64:
                r = 2
                while r <= min(lanes, max(2, size // 2)):
                    out.append(dict(size=size, per_lane=m, lanes=lanes, rest=r))
                    r *= 2
            m *= 2
    else:
        # workgroup items of at least 2 waves: 16 unknowns per lane from 8192
        # on, 8 or 16 below (the 16-per-lane form from 2048 on)
        for m in ((16,) if size >= 8192 else (8, 16)):
            lanes = size // m
            if lanes < 128:
                continue
            w = lanes // 64
            for r in sorted({2, max(2, w // 2), w, 2 * w, 4 * w, 8 * w, 16 * w, 32 * w, 64 * w}):
                if 2 <= r <= lanes:
                    out.append(dict(size=size, per_lane=m, lanes=lanes, rest=r))
    return out


def tune(sizes=None, unknowns: int = 1 << 22, iters: int = 2, reps: int = 5, log=print,
         fp64: bool = False):
    import torch

    from .ragged import RaggedPlan, offsets_of
    from .solvers import generate

    sizes = sizes or [1 << k for k in range(1, 15 if not fp64 else 14)]
    if fp64:
        iters = 0
    table = []
    for size in sizes:
        N = max(64, unknowns // size)
        offs = offsets_of([size] * N)
        arrs = [torch.empty(N * size, dtype=torch.float64, device="cuda") for _ in range(5)]
        generate(0, 7, *arrs[:4], N * size, 1)
        best = None
        for cand in candidates(size):
            plan = RaggedPlan(offs, table=[cand])
            plan.solve(*arrs, iters=iters, graph=False, fp64=fp64)
            torch.cuda.synchronize()
            times = []
            for _ in range(reps):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                plan.solve(*arrs, iters=iters, graph=False, fp64=fp64)
                e1.record()
                torch.cuda.synchronize()
                times.append(e0.elapsed_time(e1))
            plan.close()
            ms = sorted(times)[len(times) // 2]
            if best is None or ms < best["ms"]:
                best = dict(cand, ms=ms, unknowns_per_s=N * size / (ms * 1e-3))
        log(f"size {size}: {best}")
        table.append(best)
    return table


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mode", default="refine", choices=["refine", "fp64"])
    ap.add_argument("--out", default=None)
    ap.add_argument("--unknowns", type=int, default=1 << 22)
    ap.add_argument("--iters", type=int, default=2)
    args = ap.parse_args()
    import torch
    fp64 = args.mode == "fp64"
    out = args.out or os.path.join(os.path.dirname(__file__), "tune",
                                   f"ragged_f64_{args.mode}.json")
    t0 = time.time()
    table = tune(unknowns=args.unknowns, iters=args.iters, fp64=fp64)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    args.out = out
    with open(args.out, "w") as f:
        json.dump({"device": torch.cuda.get_device_name(0), "mode": args.mode,
                   "iters": 0 if fp64 else args.iters,
                   "unknowns_per_bucket": args.unknowns, "seconds": round(time.time() - t0, 1),
                   "buckets": table}, f, indent=1)
    print(args.out)


if __name__ == "__main__":
    main()
