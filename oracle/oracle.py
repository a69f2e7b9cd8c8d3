"""CPU oracle for the batched tridiagonal solve -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker. The shipped library
(``tridiagonalsolvers_amd``) never imports it.

Two independent restatements live here:

* ``thomas_*`` / ``compare`` / ``residual`` call the C oracle
  (``oracle/trid_oracle.c``, built by ``oracle/Makefile``): FP64 serial Thomas
  following ``tridiagonal_gold.h:26-44`` with OpenMP across all systems, the
  reference RMS + NaN rule of ``tridiagonal_gold.h:97-148`` and the relative
  residual.
* ``generate_np`` / ``thomas_np`` are numpy restatements (vectorised across
  systems) of the same generator and recurrence, used to cross-check the C
  code bit-for-bit (generator) and to round-off (Thomas).

Citations are relative to the reference repository's ``tridiagonalsolvers/``
directory.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

GEN_DOMINANT, GEN_WEAK, GEN_REFGEN, GEN_PAIRED, GEN_GENERAL = 0, 1, 2, 3, 4
GEN_KINDS = {"dominant": GEN_DOMINANT, "weak": GEN_WEAK, "refgen": GEN_REFGEN,
             "paired": GEN_PAIRED, "general": GEN_GENERAL}

_lib = None


def build(force: bool = False) -> str:
    """Compile the C oracle with its Makefile (gcc + OpenMP) when it is
    missing or older than its source."""
    src = os.path.join(HERE, "trid_oracle.c")
    stale = (not os.path.exists(LIB_PATH) or
             (os.path.exists(src) and os.path.getmtime(src) > os.path.getmtime(LIB_PATH)))
    if force or stale:
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"],
                       check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        for name in ("oracle_thomas_f64_from_f32", "oracle_thomas_f64",
                     "oracle_thomas_f32"):
            f = getattr(L, name)
            f.argtypes = [P, P, P, P, P, i64, i64, ctypes.c_int]
            f.restype = ctypes.c_int
        for name in ("oracle_thomas_ragged_f64_from_f32",
                     "oracle_thomas_ragged_f64"):
            f = getattr(L, name)
            f.argtypes = [P, P, P, P, P, P, i64, ctypes.c_int]
            f.restype = ctypes.c_int
        L.oracle_compare_ragged.argtypes = [P, P, P, i64, ctypes.c_double, P, P]
        L.oracle_compare_ragged.restype = i64
        L.oracle_residual_ragged.argtypes = [P, P, P, P, P, P, i64, P]
        L.oracle_residual_ragged.restype = None
        for name in ("oracle_generate_f32", "oracle_generate_f64"):
            f = getattr(L, name)
            f.argtypes = [ctypes.c_int, ctypes.c_uint64, P, P, P, P, P, i64]
            f.restype = ctypes.c_int
        for name in ("oracle_gtsv_f32", "oracle_gtsv_f64"):
            f = getattr(L, name)
            f.argtypes = [P, P, P, P, P, i64, i64, P, ctypes.c_int]
            f.restype = ctypes.c_int
        L.oracle_stream_key.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_stream_key.restype = ctypes.c_uint64
        L.oracle_u01.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_u01.restype = ctypes.c_double
        L.oracle_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(arr: np.ndarray) -> ctypes.c_void_p:
    assert arr.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(arr.ctypes.data)


def uniform_offsets(n: int, num_systems: int) -> np.ndarray:
    return np.arange(num_systems + 1, dtype=np.int64) * int(n)


def max_threads() -> int:
    return int(lib().oracle_max_threads())


# --------------------------------------------------------------------------
# generator (C oracle)
# --------------------------------------------------------------------------

def generate(kind, seed: int, n: int, num_systems: int, dtype=np.float32,
             offsets: np.ndarray | None = None):
    """Seeded diagonally dominant batch, flat system-major arrays a,b,c,d."""
    k = GEN_KINDS[kind] if isinstance(kind, str) else int(kind)
    if offsets is None:
        offsets = uniform_offsets(n, num_systems)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    num_systems = len(offsets) - 1
    total = int(offsets[-1])
    dtype = np.dtype(dtype)
    a, b, c, d = (np.empty(total, dtype=dtype) for _ in range(4))
    f = lib().oracle_generate_f32 if dtype == np.float32 else lib().oracle_generate_f64
    rc = f(k, ctypes.c_uint64(seed), _p(a), _p(b), _p(c), _p(d), _p(offsets),
           num_systems)
    if rc != 0:
        raise ValueError("oracle generator rejected its arguments")
    return a, b, c, d


# --------------------------------------------------------------------------
# generator (independent numpy restatement)
# --------------------------------------------------------------------------

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = z + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def stream_key_np(seed: int, stream: int) -> np.uint64:
    with np.errstate(over="ignore"):
        s = _mix64(np.array([stream + 1], dtype=np.uint64))
        return _mix64(np.array([seed], dtype=np.uint64) ^ s)[0]


def u01_np(seed: int, stream: int, g: np.ndarray) -> np.ndarray:
    key = stream_key_np(seed, stream)
    with np.errstate(over="ignore"):
        h = _mix64(np.asarray(g, dtype=np.uint64) ^ key)
    return (h >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def generate_np(kind, seed: int, n: int, num_systems: int, dtype=np.float32,
                offsets: np.ndarray | None = None):
    k = GEN_KINDS[kind] if isinstance(kind, str) else int(kind)
    if offsets is None:
        offsets = uniform_offsets(n, num_systems)
    offsets = np.asarray(offsets, dtype=np.int64)
    total = int(offsets[-1])
    g = np.arange(total, dtype=np.uint64)
    ua, ub, uc, ud = (u01_np(seed, s, g) for s in range(4))
    first = offsets[:-1]
    last = offsets[1:] - 1
    dt = np.dtype(dtype)
    if k == GEN_PAIRED:
        j = g.astype(np.int64) - np.repeat(first, np.diff(offsets))
        even = (j & 1) == 0
        is_last = np.zeros(total, dtype=bool)
        is_last[last] = True
        a = np.where(even, 0.25 * (2.0 * ua - 1.0), 1.0 + ua).astype(dt)
        b = np.where(even & is_last, 1.0 + ub, 0.001 * (2.0 * ub - 1.0)).astype(dt)
        c = np.where(even, 1.0 + uc, 0.25 * (2.0 * uc - 1.0)).astype(dt)
        d = (2.0 * ud - 1.0).astype(dt)
        a[first] = 0
        c[last] = 0
        return a, b, c, d
    if k == GEN_GENERAL:
        a, b, c, d = ((2.0 * u - 1.0).astype(dt) for u in (ua, ub, uc, ud))
        a[first] = 0
        c[last] = 0
        return a, b, c, d
    if k == GEN_REFGEN:
        a = (3.0 + ua).astype(dt)
        b = (8.0 + ub).astype(dt)
        c = (2.0 + uc).astype(dt)
        d = (5.0 + ud).astype(dt)
        a[first] = 0
        c[last] = 0
        return a, b, c, d
    a = (2.0 * ua - 1.0).astype(dt)
    c = (2.0 * uc - 1.0).astype(dt)
    d = (2.0 * ud - 1.0).astype(dt)
    a[first] = 0
    c[last] = 0
    margin = (1.0 + ub) if k == GEN_DOMINANT else (0.01 + 0.09 * ub)
    b = ((np.abs(a.astype(np.float64)) + np.abs(c.astype(np.float64))) + margin).astype(dt)
    return a, b, c, d


# --------------------------------------------------------------------------
# Thomas (C oracle)
# --------------------------------------------------------------------------

def thomas(a, b, c, d, n: int | None = None, offsets: np.ndarray | None = None,
           threads: int = 0) -> np.ndarray:
    """FP64 Thomas on every system (uniform size n, or ragged offsets)."""
    a, b, c, d = (np.ascontiguousarray(v) for v in (a, b, c, d))
    dt = a.dtype
    assert b.dtype == dt and c.dtype == dt and d.dtype == dt
    x = np.empty(a.shape[0], dtype=np.float64)
    L = lib()
    if offsets is None:
        assert n is not None and a.shape[0] % n == 0
        f = L.oracle_thomas_f64_from_f32 if dt == np.float32 else L.oracle_thomas_f64
        rc = f(_p(a), _p(b), _p(c), _p(d), _p(x), n, a.shape[0] // n, threads)
    else:
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        f = (L.oracle_thomas_ragged_f64_from_f32 if dt == np.float32
             else L.oracle_thomas_ragged_f64)
        rc = f(_p(a), _p(b), _p(c), _p(d), _p(x), _p(offsets),
               len(offsets) - 1, threads)
    if rc != 0:
        raise RuntimeError("oracle Thomas failed")
    return x


def thomas_f32(a, b, c, d, n: int, threads: int = 0) -> np.ndarray:
    """Same recurrence in FP32 arithmetic (the like-for-like CPU baseline)."""
    a, b, c, d = (np.ascontiguousarray(v, dtype=np.float32) for v in (a, b, c, d))
    x = np.empty_like(a)
    rc = lib().oracle_thomas_f32(_p(a), _p(b), _p(c), _p(d), _p(x), n,
                                 a.shape[0] // n, threads)
    if rc != 0:
        raise RuntimeError("oracle Thomas failed")
    return x


def gtsv(a, b, c, d, n: int, threads: int = 0):
    """Partial-pivoting solve (LAPACK ?gtsv restated, in the arrays' own
    precision: float32 = sgtsv, float64 = dgtsv) of every system of a uniform
    batch. The algorithm of the reference's MKL baseline,
    tridiagonalMKL.cpp:5-26. Returns (x, info); a system with info != 0 has
    x = NaN."""
    a, b, c, d = (np.ascontiguousarray(v) for v in (a, b, c, d))
    dt = a.dtype
    assert dt in (np.float32, np.float64)
    assert b.dtype == dt and c.dtype == dt and d.dtype == dt
    assert a.shape[0] % n == 0
    N = a.shape[0] // n
    x = np.empty_like(a)
    info = np.empty(N, dtype=np.int32)
    f = lib().oracle_gtsv_f32 if dt == np.float32 else lib().oracle_gtsv_f64
    if f(_p(a), _p(b), _p(c), _p(d), _p(x), n, N, _p(info), threads) != 0:
        raise RuntimeError("oracle gtsv failed")
    return x, info


def thomas_np(a, b, c, d, n: int) -> np.ndarray:
    """numpy restatement of tridiagonal_gold.h:26-44 in FP64, vectorised over
    systems (loop over the equation index)."""
    A, B, C, D = (np.asarray(v, dtype=np.float64).reshape(-1, n) for v in (a, b, c, d))
    N = A.shape[0]
    cp = np.empty((N, n))
    dp = np.empty((N, n))
    C = C.copy()
    C[:, n - 1] = 0.0
    cp[:, 0] = C[:, 0] / B[:, 0]
    dp[:, 0] = D[:, 0] / B[:, 0]
    for i in range(1, n):
        m = B[:, i] - A[:, i] * cp[:, i - 1]
        cp[:, i] = C[:, i] / m
        dp[:, i] = (D[:, i] - dp[:, i - 1] * A[:, i]) / m
    x = np.empty((N, n))
    x[:, n - 1] = dp[:, n - 1]
    for i in range(n - 2, -1, -1):
        x[:, i] = dp[:, i] - cp[:, i] * x[:, i + 1]
    return x.reshape(-1)


# --------------------------------------------------------------------------
# metrics
# --------------------------------------------------------------------------

def compare(x, xref, n: int | None = None, offsets=None, rms_eps: float = 1e-3):
    """Per-system (relerr, rms) and the reference failure count
    (RMS > rms_eps or NaN, tridiagonal_gold.h:131-148)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    xref = np.ascontiguousarray(xref, dtype=np.float64)
    if offsets is None:
        offsets = uniform_offsets(n, x.shape[0] // n)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    ns = len(offsets) - 1
    rel = np.empty(ns)
    rms = np.empty(ns)
    fails = lib().oracle_compare_ragged(_p(x), _p(xref), _p(offsets), ns,
                                        float(rms_eps), _p(rel), _p(rms))
    return rel, rms, int(fails)


def residual(a, b, c, d, x, n: int | None = None, offsets=None) -> np.ndarray:
    """Per-system ||d - T x||_inf / ||d||_inf in long double."""
    arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c, d, x)]
    if offsets is None:
        offsets = uniform_offsets(n, arrs[0].shape[0] // n)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    ns = len(offsets) - 1
    out = np.empty(ns)
    lib().oracle_residual_ragged(*(_p(v) for v in arrs), _p(offsets), ns, _p(out))
    return out


# --------------------------------------------------------------------------
# ADI (2D heat equation, Peaceman-Rachford), FP64, for checking the device
# ADI steps: each half step solves n tridiag(-r, 1+2r, -r) systems with the
# oracle's Thomas.
# --------------------------------------------------------------------------

def _sweep_rows(rhs: np.ndarray, r: float) -> np.ndarray:
    n = rhs.shape[1]
    m = rhs.shape[0]
    a = np.full(m * n, -r)
    c = np.full(m * n, -r)
    b = np.full(m * n, 1.0 + 2.0 * r)
    return thomas(a, b, c, np.ascontiguousarray(rhs).reshape(-1), n=n).reshape(m, n)


def adi_step_np(u: np.ndarray, r: float) -> np.ndarray:
    """(I - r dxx) u* = (I + r dyy) u; (I - r dyy) u' = (I + r dxx) u*,
    zero Dirichlet boundary; rows are x (contiguous), columns y."""
    u = np.asarray(u, dtype=np.float64)
    up = np.zeros_like(u)
    dn = np.zeros_like(u)
    up[1:] = u[:-1]
    dn[:-1] = u[1:]
    us = _sweep_rows(u + r * (up - 2 * u + dn), r)
    lf = np.zeros_like(us)
    rt = np.zeros_like(us)
    lf[:, 1:] = us[:, :-1]
    rt[:, :-1] = us[:, 1:]
    rhs = us + r * (lf - 2 * us + rt)
    return _sweep_rows(rhs.T, r).T


def adi_mode_factor(n: int, p: int, q: int, r: float) -> float:
    """Exact one-step amplification of the mode sin(p pi i/(n+1)) sin(q pi j/(n+1))."""
    lp = 4 * np.sin(p * np.pi / (2 * (n + 1))) ** 2
    lq = 4 * np.sin(q * np.pi / (2 * (n + 1))) ** 2
    return (1 - r * lp) * (1 - r * lq) / ((1 + r * lp) * (1 + r * lq))
