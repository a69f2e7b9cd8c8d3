"""Python client of the continuous-batching request server (trid_server_* in
include/trid.h). The server itself is C++ (trid_server.cpp): a scheduler
thread admits requests by device-memory footprint, batches arrivals into
ragged solves (FP32 + FP64 refinement) on several streams and copies results
back. This wrapper keeps each request's numpy arrays alive until it is
waited for.
"""
from __future__ import annotations

import ctypes
import threading

import numpy as np

from ._lib import (EAGAIN, REQ_ADI, REQ_OPERATOR, REQ_SOLVE, BucketCfg, Request,
                   RequestResult, ServerOpts, ServerStats, TridError, check, lib)
from .ragged import load_table

FAKE_EXECUTOR = 1
STATES = {2: "done", 3: "cancelled", 4: "failed"}


def _f64(v) -> np.ndarray:
    v = np.ascontiguousarray(v, dtype=np.float64)
    return v


class Server:
    def __init__(self, hbm_budget_bytes: int = 4 << 30, slots: int = 3,
                 max_batch_systems: int = 65536, max_batch_unknowns: int = 1 << 23,
                 batch_window_us: int = 200, iters: int = 2, fake: bool = False,
                 fake_fail_every: int = 0, fake_batch_us: int = 0, table="default"):
        if table == "default":
            table = load_table()
        table = table or []
        self._table = (BucketCfg * max(1, len(table)))()
        for i, t in enumerate(table):
            self._table[i] = BucketCfg(t["size"], t["per_lane"], t["lanes"], t["rest"])
        o = ServerOpts(int(hbm_budget_bytes), int(slots), int(max_batch_systems),
                       int(max_batch_unknowns), int(batch_window_us), int(iters),
                       FAKE_EXECUTOR if fake else 0, int(fake_fail_every), int(fake_batch_us),
                       ctypes.cast(self._table, ctypes.c_void_p), len(table))
        h = ctypes.c_void_p()
        check(lib().trid_server_create(ctypes.byref(h), ctypes.byref(o)), "trid_server_create")
        self._h = h
        self._keep = {}
        self._ops = {}
        self._host = {}
        self._mu = threading.Lock()

    # -- page-locked request memory ---------------------------------------------
    def host_array(self, n: int) -> np.ndarray:
        """n float64 values of page-locked memory owned by the server (valid
        until host_free or close). Request arrays placed inside such memory
        go to and from the device by DMA with no staging copy."""
        p = ctypes.c_void_p()
        check(lib().trid_server_host_alloc(self._h, int(n) * 8, ctypes.byref(p)),
              "trid_server_host_alloc")
        arr = np.ctypeslib.as_array((ctypes.c_double * int(n)).from_address(p.value))
        self._host[arr.ctypes.data] = p.value
        return arr

    def host_free(self, arr: np.ndarray) -> None:
        p = self._host.pop(arr.ctypes.data, None)
        if p is None:
            raise ValueError("not an array from host_array")
        check(lib().trid_server_host_free(self._h, ctypes.c_void_p(p)), "trid_server_host_free")

    # -- operators ---------------------------------------------------------
    def register_operator(self, sizes, a, b, c) -> int:
        sizes = np.ascontiguousarray(sizes, dtype=np.int64)
        a, b, c = _f64(a), _f64(b), _f64(c)
        op = ctypes.c_int32()
        check(lib().trid_server_register_operator(
            self._h, ctypes.c_void_p(sizes.ctypes.data), len(sizes),
            *(ctypes.c_void_p(v.ctypes.data) for v in (a, b, c)), ctypes.byref(op)),
            "register_operator")
        self._ops[op.value] = int(sizes.sum())
        return op.value

    def release_operator(self, op: int) -> int:
        return lib().trid_server_release_operator(self._h, int(op))

    # -- requests ----------------------------------------------------------
    def submit(self, d, x, offsets=None, a=None, b=None, c=None, op: int = -1) -> int:
        """Queue a request; returns its ticket. d, x (and a, b, c unless op is
        given) are host float64 arrays; offsets has num_systems+1 entries."""
        d = _f64(d)
        if x.dtype != np.float64 or not x.flags["C_CONTIGUOUS"]:
            raise TypeError("x must be a C-contiguous float64 array")
        keep = [d, x]
        r = Request()
        r.kind = REQ_OPERATOR if op >= 0 else REQ_SOLVE
        r.op = int(op)
        r.d = d.ctypes.data
        r.x = x.ctypes.data
        if op < 0:
            offsets = np.ascontiguousarray(offsets, dtype=np.int64)
            a, b, c = _f64(a), _f64(b), _f64(c)
            keep += [offsets, a, b, c]
            r.num_systems = len(offsets) - 1
            r.offsets = offsets.ctypes.data
            r.a, r.b, r.c = a.ctypes.data, b.ctypes.data, c.ctypes.data
        t = ctypes.c_uint64()
        check(lib().trid_server_submit(self._h, ctypes.byref(r), ctypes.byref(t)), "submit")
        with self._mu:
            self._keep[t.value] = keep
        return t.value

    # -- resident ADI grids ------------------------------------------------
    def create_adi(self, u0) -> int:
        u0 = _f64(u0)
        n = int(round(u0.size ** 0.5))
        if n * n != u0.size:
            raise ValueError("u0 must hold an n x n grid")
        adi = ctypes.c_int32()
        check(lib().trid_server_create_adi(self._h, n, ctypes.c_void_p(u0.ctypes.data),
                                           ctypes.byref(adi)), "create_adi")
        return adi.value

    def submit_adi(self, adi: int, r: float, steps: int = 1) -> int:
        q = Request()
        q.kind = REQ_ADI
        q.adi = int(adi)
        q.adi_steps = int(steps)
        q.adi_r = float(r)
        t = ctypes.c_uint64()
        check(lib().trid_server_submit(self._h, ctypes.byref(q), ctypes.byref(t)), "submit_adi")
        return t.value

    def read_adi(self, adi: int, n: int) -> np.ndarray:
        u = np.empty(n * n, dtype=np.float64)
        check(lib().trid_server_read_adi(self._h, int(adi), ctypes.c_void_p(u.ctypes.data)),
              "read_adi")
        return u.reshape(n, n)

    def release_adi(self, adi: int) -> int:
        return lib().trid_server_release_adi(self._h, int(adi))

    def wait(self, ticket: int, timeout_us: int = -1) -> dict | None:
        res = RequestResult()
        st = lib().trid_server_wait(self._h, ctypes.c_uint64(ticket), int(timeout_us),
                                    ctypes.byref(res))
        if st == EAGAIN:
            return None
        check(st, "wait")
        with self._mu:
            self._keep.pop(ticket, None)
        out = {k: getattr(res, k) for k, _ in RequestResult._fields_}
        out["state"] = STATES.get(out["state"], out["state"])
        return out

    def cancel(self, ticket: int) -> bool:
        return lib().trid_server_cancel(self._h, ctypes.c_uint64(ticket)) == 0

    def stats(self) -> dict:
        s = ServerStats()
        check(lib().trid_server_get_stats(self._h, ctypes.byref(s)), "stats")
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def close(self):
        if self._h:
            lib().trid_server_destroy(self._h)   # frees the host arrays too
            self._h = None
            self._keep.clear()
            self._host.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["Server", "TridError", "FAKE_EXECUTOR"]
