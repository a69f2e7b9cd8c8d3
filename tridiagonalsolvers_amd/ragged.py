"""Ragged batches: one persistent, graph-captured launch (trid_ragged_* in
include/trid.h), the per-bucket autotune table it reads, and helpers.

System s of a ragged batch occupies [offsets[s], offsets[s+1]) of flat FP64
device arrays. A RaggedPlan depends only on the sizes and the bucket table.
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np

from ._lib import BucketCfg, RaggedInfo, check, lib
from .solvers import _dev_ptrs, _ptr, _stream_ptr

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(HERE, "tune", "ragged_f64_refine.json")
# bucket table tuned for the one-pass FP64 mode (RaggedPlan(..., "default-fp64"))
DEFAULT_TABLE_FP64 = os.path.join(HERE, "tune", "ragged_f64_fp64.json")
NO_GRAPH = 1
FP64 = 2


def log_uniform_sizes(num_systems: int, lo: int, hi: int, seed: int) -> np.ndarray:
    """Sizes round(exp(U[ln lo, ln hi])), seeded (numpy PCG64)."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(np.log(lo), np.log(hi), size=num_systems)
    return np.clip(np.rint(np.exp(u)).astype(np.int64), lo, hi)


def offsets_of(sizes) -> np.ndarray:
    return np.concatenate([[0], np.cumsum(np.asarray(sizes, dtype=np.int64))]).astype(np.int64)


def load_table(path: str | None = None) -> list[dict]:
    """Bucket configurations written by tridiagonalsolvers_amd.autotune."""
    path = DEFAULT_TABLE if path is None else path
    if not os.path.exists(path):
        return []
    with open(path) as f:
        return [{k: int(b[k]) for k in ("size", "per_lane", "lanes", "rest")}
                for b in json.load(f)["buckets"]]


class RaggedPlan:
    def __init__(self, offsets, table: list[dict] | None | str = "default"):
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.offsets = offsets
        self.num_systems = len(offsets) - 1
        if table == "default":
            table = load_table()
        elif table == "default-fp64":
            table = load_table(DEFAULT_TABLE_FP64)
        table = table or []
        arr = (BucketCfg * max(1, len(table)))()
        for i, t in enumerate(table):
            arr[i] = BucketCfg(t["size"], t["per_lane"], t["lanes"], t["rest"])
        h = ctypes.c_void_p()
        st = lib().trid_ragged_plan_create(ctypes.byref(h), ctypes.c_void_p(offsets.ctypes.data),
                                           self.num_systems, ctypes.cast(arr, ctypes.c_void_p),
                                           len(table))
        check(st, "trid_ragged_plan_create")
        self._h = h

    def info(self) -> dict:
        inf = RaggedInfo()
        cap = 16
        bk = (BucketCfg * cap)()
        cnt = (ctypes.c_int64 * cap)()
        check(lib().trid_ragged_plan_info(self._h, ctypes.byref(inf),
                                          ctypes.cast(bk, ctypes.c_void_p),
                                          ctypes.cast(cnt, ctypes.c_void_p), cap), "plan_info")
        out = {k: getattr(inf, k) for k, _ in RaggedInfo._fields_}
        out["buckets"] = [dict(size=bk[i].size, per_lane=bk[i].per_lane, lanes=bk[i].lanes,
                               rest=bk[i].rest, systems=cnt[i])
                          for i in range(inf.num_buckets)]
        return out

    def solve(self, a, b, c, d, x, iters: int | None = None, stream=None, graph: bool = True,
              fp64: bool = False, tol: float = 0.0) -> None:
        """FP32 solve + `iters` FP64 refinements (default 2), or with fp64=True
        one FP64 solve (iters must then be 0 or None; systems up to 8192).
        tol > 0: a system stops refining once its relative residual
        max|d - A x| / max|d| is at most tol (`iters` is then the most it
        takes; trid_ragged_solve_tol_f64)."""
        if iters is None:
            iters = 0 if fp64 else 2
        flags = (0 if graph else NO_GRAPH) | (FP64 if fp64 else 0)
        ptrs = _dev_ptrs((a, b, c, d, x), int(self.offsets[-1]), "f64", "ragged solve")
        if tol:
            st = lib().trid_ragged_solve_tol_f64(self._h, *ptrs, int(iters), float(tol), flags,
                                                 ctypes.c_void_p(_stream_ptr(stream)))
            check(st, "trid_ragged_solve_tol_f64")
            return
        st = lib().trid_ragged_solve_f64(self._h, *ptrs, int(iters), flags,
                                         ctypes.c_void_p(_stream_ptr(stream)))
        check(st, "trid_ragged_solve_f64")

    def close(self):
        if self._h:
            lib().trid_ragged_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def residual(a, b, c, d, x, d_offsets, num_systems: int, out, stream=None) -> None:
    """Per-system relative residual in FP64 on the device (trid_residual_f64)."""
    offs = _dev_ptrs((d_offsets,), int(num_systems) + 1, "i64", "residual offsets")[0]
    res = _dev_ptrs((out,), int(num_systems), "f64", "residual out")[0]
    # array lengths come from the offsets on the device; dtype and placement
    # are checked here
    ptrs = _dev_ptrs((a, b, c, d, x), 0, "f64", "residual")
    st = lib().trid_residual_f64(*ptrs, offs, int(num_systems), res,
                                 ctypes.c_void_p(_stream_ptr(stream)))
    check(st, "trid_residual_f64")
