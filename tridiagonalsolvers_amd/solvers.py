"""Host-side mirror of the reference solver API, over the C ABI of libtrid.so.

The reference exposes seven free functions with host pointers, a
compile-time element type and a void return (forward-declared in
test_tridiagonal.cpp:17-26). This module keeps their names, argument order
and meaning -- ``cr(a, b, c, d, x, systemSize, numSystems)`` and so on --
for numpy arrays (float32 or float64 picks the _f32 or _f64 entry point).
``x`` is overwritten; ``a, b, c, d`` are not modified. Errors raise
``TridError`` instead of exiting the process.

``Handle.solve`` is the stream-ordered device API (``trid_solve_*``) for
device arrays: torch tensors on a GPU (checked for device, dtype,
contiguity and length before any kernel runs) or integer device addresses
(FP32 unless ``dtype="f64"`` is given).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import METHODS, Opts, PlanInfo, TridError, check, lib


def _suffix(arr: np.ndarray) -> str:
    if arr.dtype == np.float32:
        return "f32"
    if arr.dtype == np.float64:
        return "f64"
    raise TypeError(f"unsupported dtype {arr.dtype}; use float32 or float64")


def _host_ptrs(a, b, c, d, x, systemSize: int, numSystems: int):
    arrs = (a, b, c, d, x)
    for v in arrs:
        if not isinstance(v, np.ndarray):
            raise TypeError("host arrays must be numpy arrays")
        if not v.flags["C_CONTIGUOUS"]:
            raise ValueError("host arrays must be C-contiguous")
        if v.dtype != a.dtype:
            raise TypeError("a, b, c, d, x must share one dtype")
        if v.size < int(systemSize) * int(numSystems):
            raise ValueError("array shorter than systemSize * numSystems")
    if not x.flags["WRITEABLE"]:
        raise ValueError("x must be writeable")
    return [ctypes.c_void_p(v.ctypes.data) for v in arrs]


def _call7(name: str, a, b, c, d, x, systemSize, numSystems):
    p = _host_ptrs(a, b, c, d, x, systemSize, numSystems)
    f = getattr(lib(), f"trid_{name}_{_suffix(a)}")
    check(f(*p, int(systemSize), int(numSystems)), name)


def _call8(name: str, a, b, c, d, x, systemSize, numSystems, extra):
    p = _host_ptrs(a, b, c, d, x, systemSize, numSystems)
    f = getattr(lib(), f"trid_{name}_{_suffix(a)}")
    check(f(*p, int(systemSize), int(numSystems), int(extra)), name)


# --- reference names (cr_app.cu:4, pcr_app.cu:5, crpcr_app.cu:5, crNBC.cu:123,
#     crNBC.cu:344, crNBC32banks.cu:123, pcrGauss_app.cu:164) -----------------

def cr(a, b, c, d, x, systemSize, numSystems):
    _call7("cr", a, b, c, d, x, systemSize, numSystems)


def pcr(a, b, c, d, x, systemSize, numSystems):
    _call7("pcr", a, b, c, d, x, systemSize, numSystems)


def crpcr(a, b, c, d, x, systemSize, numSystems, restSystemSize):
    _call8("crpcr", a, b, c, d, x, systemSize, numSystems, restSystemSize)


def crNBC(a, b, c, d, x, systemSize, numSystems):
    _call7("crnbc", a, b, c, d, x, systemSize, numSystems)


def crpcrNBC(a, b, c, d, x, systemSize, numSystems, restSystemSize):
    _call8("crpcrnbc", a, b, c, d, x, systemSize, numSystems, restSystemSize)


def crNBC32(a, b, c, d, x, systemSize, numSystems):
    _call7("crnbc32", a, b, c, d, x, systemSize, numSystems)


def pcrGaussWrapper(a, b, c, d, x, systemSize, numSystems, gaussian):
    _call8("pcrgauss", a, b, c, d, x, systemSize, numSystems, gaussian)


def gtsv(a, b, c, d, x, systemSize, numSystems):
    """Partial-pivoting solve (LAPACK ?gtsv semantics) of every system; the
    counterpart of the reference's MKL baseline ``tridiagonalMKL``
    (tridiagonalMKL.cpp:5), which solves in place and swaps the d/x
    pointers. Here a, b, c, d are left untouched and x receives the
    solution; a singular system's x is NaN."""
    _call7("gtsv", a, b, c, d, x, systemSize, numSystems)


tridiagonalMKL = gtsv


# --- options, plans, device API ----------------------------------------------

def make_opts(rest=0, gaussian=0, lds_pad=0, lanes=0, per_lane=0, io_mode=0,
              flags=0) -> Opts:
    return Opts(int(rest), int(gaussian), int(lds_pad), int(lanes), int(per_lane),
                int(io_mode), int(flags))


def method_id(method) -> int:
    return METHODS[method] if isinstance(method, str) else int(method)


def plan_query(method, systemSize: int, numSystems: int, elem_size: int = 4,
               opts: Opts | None = None) -> dict:
    """Validate a request and report the kernel plan (no device needed)."""
    info = PlanInfo()
    o = opts if opts is not None else Opts()
    st = lib().trid_plan_query(method_id(method), int(systemSize), int(numSystems),
                               int(elem_size), ctypes.byref(o), ctypes.byref(info))
    check(st, "plan_query")
    return {k: getattr(info, k) for k, _ in PlanInfo._fields_}


def _ptr(v) -> int:
    if isinstance(v, int):
        return v
    if hasattr(v, "data_ptr"):
        return int(v.data_ptr())
    raise TypeError("device arrays must be torch tensors or integer addresses")


_TORCH_DTYPES = {"torch.float32": "f32", "torch.float64": "f64", "torch.int32": "i32",
                 "torch.int64": "i64"}


def _dtype_of(v) -> str | None:
    """'f32' / 'f64' / 'i32' / 'i64' for a torch tensor, None for a raw address."""
    if isinstance(v, int) or v is None:
        return None
    name = str(getattr(v, "dtype", ""))
    if name not in _TORCH_DTYPES:
        raise TypeError(f"unsupported device array dtype {name or type(v).__name__}")
    return _TORCH_DTYPES[name]


def _infer_dtype(a, dtype: str | None) -> str:
    if dtype is not None:
        if dtype not in ("f32", "f64"):
            raise ValueError("dtype must be 'f32' or 'f64'")
        return dtype
    t = _dtype_of(a)
    if t is None:            # integer device address: FP32 unless dtype says otherwise
        return "f32"
    if t not in ("f32", "f64"):
        raise TypeError("device arrays must be float32 or float64")
    return t


def _dev_ptrs(arrs, need: int, dtype: str, what: str = "array"):
    """Device addresses of `arrs` after checking what a torch tensor can tell:
    on a GPU, of `dtype`, contiguous and holding at least `need` elements.
    Integer addresses pass unchecked (the caller vouches for them). Raises
    before the C ABI is called, so no kernel sees a host pointer or reads past
    an array's end."""
    tensors = [v for v in arrs if _dtype_of(v) is not None]
    for v in tensors:
        if _dtype_of(v) != dtype:
            raise TypeError(f"{what}: tensor dtype {v.dtype} does not match {dtype}")
    out = []
    for v in arrs:
        if _dtype_of(v) is not None:
            if not getattr(v, "is_cuda", False):
                raise ValueError(f"{what}: tensor is not on a GPU device")
            if not v.is_contiguous():
                raise ValueError(f"{what}: tensor is not contiguous")
            if v.numel() < need:
                raise ValueError(f"{what}: tensor holds {v.numel()} elements, {need} needed")
        out.append(ctypes.c_void_p(_ptr(v)))
    return out


def _stream_ptr(stream) -> int | None:
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return int(stream.cuda_stream)


class Handle:
    """Owns a trid_handle (per-thread workspace) for the device API."""

    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().trid_create(ctypes.byref(h)), "trid_create")
        self._h = h

    def close(self):
        if self._h:
            lib().trid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, method, a, b, c, d, x, systemSize: int, numSystems: int,
              opts: Opts | None = None, stream=None, dtype: str | None = None) -> None:
        dtype = _infer_dtype(a, dtype)
        ptrs = _dev_ptrs((a, b, c, d, x), int(systemSize) * int(numSystems), dtype, "solve")
        f = getattr(lib(), f"trid_solve_{dtype}")
        o = opts if opts is not None else Opts()
        st = f(self._h, method_id(method), *ptrs, int(systemSize), int(numSystems),
               ctypes.byref(o), ctypes.c_void_p(_stream_ptr(stream)))
        check(st, "trid_solve")

    def solve_strided(self, method, a, b, c, d, x, systemSize: int, numSystems: int,
                      batchStride: int, opts: Opts | None = None, stream=None,
                      dtype: str | None = None) -> None:
        """Strided batch: system s at s * batchStride (>= systemSize), the
        layout of rocSPARSE's gtsv_no_pivot_strided_batch; register methods."""
        dtype = _infer_dtype(a, dtype)
        need = (int(numSystems) - 1) * int(batchStride) + int(systemSize) if numSystems > 0 else 0
        ptrs = _dev_ptrs((a, b, c, d, x), need, dtype, "solve_strided")
        f = getattr(lib(), f"trid_solve_strided_{dtype}")
        o = opts if opts is not None else Opts()
        st = f(self._h, method_id(method), *ptrs, int(systemSize), int(numSystems),
               int(batchStride), ctypes.byref(o), ctypes.c_void_p(_stream_ptr(stream)))
        check(st, "trid_solve_strided")

    def solve_interleaved(self, a, b, c, d, x, systemSize: int, numSystems: int,
                          batchStride: int | None = None, stream=None,
                          dtype: str | None = None) -> None:
        """Interleaved batch: row i of system j at i * batchStride + j
        (default batchStride = numSystems); x may be d."""
        dtype = _infer_dtype(a, dtype)
        stride = numSystems if batchStride is None else batchStride
        need = (int(systemSize) - 1) * int(stride) + int(numSystems) if numSystems > 0 else 0
        ptrs = _dev_ptrs((a, b, c, d, x), need, dtype, "solve_interleaved")
        f = getattr(lib(), f"trid_solve_interleaved_{dtype}")
        st = f(self._h, *ptrs, int(systemSize), int(numSystems), int(stride),
               ctypes.c_void_p(_stream_ptr(stream)))
        check(st, "trid_solve_interleaved")

    def solve_pivoting(self, a, b, c, d, x, systemSize: int, numSystems: int,
                       info=None, batchStride: int | None = None, stream=None,
                       dtype: str | None = None) -> None:
        """Partial pivoting (LAPACK ?gtsv). System-major by default; with
        batchStride, interleaved (row i of system j at i * batchStride + j).
        info: optional device int32 array of numSystems (0, or the 1-based
        row of the first zero pivot); x may be d."""
        dtype = _infer_dtype(a, dtype)
        if batchStride is None:
            need = int(systemSize) * int(numSystems)
        else:
            need = (int(systemSize) - 1) * int(batchStride) + int(numSystems) if numSystems > 0 else 0
        ptrs = _dev_ptrs((a, b, c, d, x), need, dtype, "solve_pivoting")
        info_p = ctypes.c_void_p(None)
        if info is not None:
            info_p = _dev_ptrs((info,), int(numSystems), "i32", "solve_pivoting info")[0]
        s = ctypes.c_void_p(_stream_ptr(stream))
        if batchStride is None:
            st = getattr(lib(), f"trid_solve_pivoting_{dtype}")(
                self._h, *ptrs, int(systemSize), int(numSystems), info_p, s)
        else:
            st = getattr(lib(), f"trid_solve_pivoting_interleaved_{dtype}")(
                self._h, *ptrs, int(systemSize), int(numSystems), int(batchStride), info_p, s)
        check(st, "trid_solve_pivoting")

    def solve_refined(self, method, a, b, c, d, x, systemSize: int, numSystems: int,
                      iters: int = 2, opts: Opts | None = None, stream=None) -> None:
        """FP64 arrays; FP32 solves plus `iters` FP64 residual corrections."""
        o = opts if opts is not None else Opts()
        ptrs = _dev_ptrs((a, b, c, d, x), int(systemSize) * int(numSystems), "f64",
                         "solve_refined")
        st = lib().trid_solve_refined_f64(
            self._h, method_id(method), *ptrs, int(systemSize), int(numSystems), int(iters),
            ctypes.byref(o), ctypes.c_void_p(_stream_ptr(stream)))
        check(st, "trid_solve_refined_f64")


def generate(kind: int, seed: int, a, b, c, d, systemSize: int, numSystems: int,
             stream=None, dtype: str | None = None) -> None:
    """Seeded synthetic batch written on the device (trid_generate_*)."""
    dtype = _infer_dtype(a, dtype)
    ptrs = _dev_ptrs((a, b, c, d), int(systemSize) * int(numSystems), dtype, "generate")
    f = getattr(lib(), f"trid_generate_{dtype}")
    st = f(int(kind), ctypes.c_uint64(seed), *ptrs, int(systemSize), int(numSystems),
           ctypes.c_void_p(_stream_ptr(stream)))
    check(st, "trid_generate")


def device_count() -> int:
    return int(lib().trid_device_count())


__all__ = ["cr", "pcr", "crpcr", "crNBC", "crpcrNBC", "crNBC32", "pcrGaussWrapper",
           "gtsv", "tridiagonalMKL",
           "Handle", "generate", "plan_query", "make_opts", "device_count", "TridError",
           "METHODS", "_lib"]
