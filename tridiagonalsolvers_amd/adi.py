"""Resident ADI grid (trid_adi_* in include/trid.h): Peaceman-Rachford steps
of the 2D heat equation on an n x n FP64 device grid. Every system of both
sweeps has the operator tridiag(-r, 1+2r, -r): each call computes its LU
factors once (closed-form pivots, resident beside the grid) and applies them
to the rows and, on the transposed grid, to the columns (kern_adi.hip's
factored-operator kernel: the two substitutions as lane scans)."""
from __future__ import annotations

import ctypes

from ._lib import check, lib
from .solvers import _ptr, _stream_ptr


class AdiGrid:
    def __init__(self, n: int, u=None):
        """u: an n*n float64 torch tensor on cuda to step in place (kept
        referenced), or None to let the library allocate the grid."""
        h = ctypes.c_void_p()
        self._u = u
        check(lib().trid_adi_create(ctypes.byref(h), int(n),
                                    ctypes.c_void_p(_ptr(u) if u is not None else None)),
              "trid_adi_create")
        self._h = h
        self.n = int(n)

    def grid_ptr(self) -> int:
        return int(lib().trid_adi_grid(self._h))

    def step(self, r: float, steps: int = 1, stream=None) -> None:
        check(lib().trid_adi_step(self._h, float(r), int(steps),
                                  ctypes.c_void_p(_stream_ptr(stream))), "trid_adi_step")

    def close(self):
        if self._h:
            lib().trid_adi_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
