"""ctypes binding of libtrid.so (the C ABI declared in include/trid.h).

The library is loaded from this package directory. There is no fallback:
if libtrid.so is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# TRID_LIBRARY selects another build of the same ABI (the LDS-poisoning
# debug library libtrid_lds_poison.so); default: the in-tree libtrid.so
LIB_PATH = os.environ.get("TRID_LIBRARY") or os.path.join(HERE, "libtrid.so")

# trid_status
OK, EINVAL, ENOMEM, EHIP, ENOTSUP, ENODEV, EAGAIN = range(7)
# trid_method
CR, PCR, CRPCR, CRNBC, CRPCRNBC, CRNBC32, PCRGAUSS, AUTO = range(8)
METHODS = {"cr": CR, "pcr": PCR, "crpcr": CRPCR, "crnbc": CRNBC, "crpcrnbc": CRPCRNBC,
           "crnbc32": CRNBC32, "pcrgauss": PCRGAUSS, "auto": AUTO}

# symbols declared in include/trid.h, checked at load time
EXPORTS = (
    "trid_status_string", "trid_last_status", "trid_device_count",
    "trid_cr_f32", "trid_cr_f64", "trid_pcr_f32", "trid_pcr_f64",
    "trid_crpcr_f32", "trid_crpcr_f64", "trid_crnbc_f32", "trid_crnbc_f64",
    "trid_crpcrnbc_f32", "trid_crpcrnbc_f64", "trid_crnbc32_f32", "trid_crnbc32_f64",
    "trid_pcrgauss_f32", "trid_pcrgauss_f64",
    "trid_create", "trid_destroy", "trid_solve_f32", "trid_solve_f64",
    "trid_generate_f32", "trid_generate_f64", "trid_plan_query",
    "trid_solve_refined_f64", "trid_residual_f64", "trid_solve_noalloc",
    "trid_solve_interleaved_f32", "trid_solve_interleaved_f64",
    "trid_solve_strided_f32", "trid_solve_strided_f64",
    "trid_solve_pivoting_f32", "trid_solve_pivoting_f64",
    "trid_solve_pivoting_interleaved_f32", "trid_solve_pivoting_interleaved_f64",
    "trid_gtsv_f32", "trid_gtsv_f64",
    "trid_ragged_plan_create", "trid_ragged_plan_destroy", "trid_ragged_plan_info",
    "trid_ragged_solve_f64", "trid_ragged_solve_tol_f64", "trid_bucket_table_load",
    "trid_server_create", "trid_server_destroy", "trid_server_register_operator",
    "trid_server_release_operator", "trid_server_submit", "trid_server_wait",
    "trid_server_cancel", "trid_server_get_stats", "trid_server_host_alloc",
    "trid_server_host_free",
    "trid_adi_create", "trid_adi_destroy", "trid_adi_grid", "trid_adi_step",
    "trid_server_create_adi", "trid_server_read_adi", "trid_server_release_adi",
)


class TridError(RuntimeError):
    def __init__(self, status: int, where: str = ""):
        self.status = int(status)
        msg = string(status) if _lib is not None else f"status {status}"
        super().__init__(f"{where}: {msg}" if where else msg)


class Opts(ctypes.Structure):
    _fields_ = [("rest", ctypes.c_int32), ("gaussian", ctypes.c_int32),
                ("lds_pad", ctypes.c_int32), ("lanes", ctypes.c_int32),
                ("per_lane", ctypes.c_int32), ("io_mode", ctypes.c_int32),
                ("flags", ctypes.c_int32)]


class BucketCfg(ctypes.Structure):
    _fields_ = [("size", ctypes.c_int32), ("per_lane", ctypes.c_int32),
                ("lanes", ctypes.c_int32), ("rest", ctypes.c_int32)]


class RaggedInfo(ctypes.Structure):
    _fields_ = [("num_systems", ctypes.c_int64), ("total_unknowns", ctypes.c_int64),
                ("num_buckets", ctypes.c_int32), ("workgroup_items", ctypes.c_int32),
                ("wave_items", ctypes.c_int32), ("grid", ctypes.c_int32)]


class ServerOpts(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("slots", ctypes.c_int32),
                ("max_batch_systems", ctypes.c_int32), ("max_batch_unknowns", ctypes.c_int64),
                ("batch_window_us", ctypes.c_int32), ("iters", ctypes.c_int32),
                ("flags", ctypes.c_int32), ("fake_fail_every", ctypes.c_int32),
                ("fake_batch_us", ctypes.c_int32),
                ("bucket_table", ctypes.c_void_p), ("bucket_table_size", ctypes.c_int32)]


REQ_SOLVE, REQ_OPERATOR, REQ_ADI = 0, 1, 2


class Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("num_systems", ctypes.c_int64),
                ("offsets", ctypes.c_void_p),
                ("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("c", ctypes.c_void_p),
                ("d", ctypes.c_void_p), ("x", ctypes.c_void_p), ("op", ctypes.c_int32),
                ("adi", ctypes.c_int32), ("adi_steps", ctypes.c_int32),
                ("adi_r", ctypes.c_double)]


class RequestResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("state", ctypes.c_int32),
                ("latency_us", ctypes.c_double), ("queue_us", ctypes.c_double),
                ("batch", ctypes.c_int64), ("batch_systems", ctypes.c_int64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64),
                ("failed", ctypes.c_int64), ("cancelled", ctypes.c_int64),
                ("batches", ctypes.c_int64), ("queued", ctypes.c_int64),
                ("device_bytes_in_use", ctypes.c_int64), ("peak_device_bytes", ctypes.c_int64),
                ("budget_bytes", ctypes.c_int64)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("family", ctypes.c_int32), ("per_lane", ctypes.c_int32),
                ("lanes", ctypes.c_int32), ("rest", ctypes.c_int32),
                ("lds_pad", ctypes.c_int32), ("gaussian", ctypes.c_int32),
                ("padded_size", ctypes.c_int64)]


_lib = None


def lib() -> ctypes.CDLL:
    """Load libtrid.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    try:
        # libtrid.so and torch both bind libamdhip64.so.7 by soname. Loading
        # torch first makes the process use one HIP runtime (torch's bundled
        # copy) for both; the other order leaves torch unable to initialise.
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() "
            "(hipcc --offload-arch=gfx950). There is no CPU fallback.")
    L = ctypes.CDLL(LIB_PATH)
    missing = [s for s in EXPORTS if not hasattr(L, s)]
    if missing:
        raise ImportError(f"libtrid.so lacks symbols: {missing}")
    P = ctypes.c_void_p
    i32, i64 = ctypes.c_int, ctypes.c_int64
    L.trid_status_string.argtypes = [i32]
    L.trid_status_string.restype = ctypes.c_char_p
    L.trid_last_status.restype = i32
    L.trid_device_count.restype = i32
    for name in ("cr", "pcr", "crnbc", "crnbc32"):
        for t in ("f32", "f64"):
            f = getattr(L, f"trid_{name}_{t}")
            f.argtypes = [P, P, P, P, P, i32, i32]
            f.restype = i32
    for t in ("f32", "f64"):
        f = getattr(L, f"trid_gtsv_{t}")
        f.argtypes = [P, P, P, P, P, i32, i32]
        f.restype = i32
        f = getattr(L, f"trid_solve_pivoting_{t}")
        f.argtypes = [P, P, P, P, P, P, i64, i64, P, P]
        f.restype = i32
        f = getattr(L, f"trid_solve_pivoting_interleaved_{t}")
        f.argtypes = [P, P, P, P, P, P, i64, i64, i64, P, P]
        f.restype = i32
    for name in ("crpcr", "crpcrnbc", "pcrgauss"):
        for t in ("f32", "f64"):
            f = getattr(L, f"trid_{name}_{t}")
            f.argtypes = [P, P, P, P, P, i32, i32, i32]
            f.restype = i32
    L.trid_create.argtypes = [ctypes.POINTER(P)]
    L.trid_create.restype = i32
    L.trid_destroy.argtypes = [P]
    L.trid_destroy.restype = i32
    for t in ("f32", "f64"):
        f = getattr(L, f"trid_solve_{t}")
        f.argtypes = [P, i32, P, P, P, P, P, i64, i64, ctypes.POINTER(Opts), P]
        f.restype = i32
        g = getattr(L, f"trid_generate_{t}")
        g.argtypes = [i32, ctypes.c_uint64, P, P, P, P, i64, i64, P]
        g.restype = i32
        v = getattr(L, f"trid_solve_interleaved_{t}")
        v.argtypes = [P, P, P, P, P, P, i64, i64, i64, P]
        v.restype = i32
        v = getattr(L, f"trid_solve_strided_{t}")
        v.argtypes = [P, i32, P, P, P, P, P, i64, i64, i64, ctypes.POINTER(Opts), P]
        v.restype = i32
    L.trid_solve_noalloc.argtypes = [i32, i32, P, P, P, P, P, i64, i64, ctypes.POINTER(Opts), P]
    L.trid_solve_noalloc.restype = i32
    L.trid_solve_refined_f64.argtypes = [P, i32, P, P, P, P, P, i64, i64, i32,
                                         ctypes.POINTER(Opts), P]
    L.trid_solve_refined_f64.restype = i32
    L.trid_residual_f64.argtypes = [P, P, P, P, P, P, i64, P, P]
    L.trid_residual_f64.restype = i32
    L.trid_ragged_plan_create.argtypes = [ctypes.POINTER(P), P, i64, P, i32]
    L.trid_ragged_plan_create.restype = i32
    L.trid_ragged_plan_destroy.argtypes = [P]
    L.trid_ragged_plan_destroy.restype = i32
    L.trid_ragged_plan_info.argtypes = [P, ctypes.POINTER(RaggedInfo), P, P, i32]
    L.trid_ragged_plan_info.restype = i32
    L.trid_bucket_table_load.argtypes = [ctypes.c_char_p, P, i32, ctypes.POINTER(i32)]
    L.trid_bucket_table_load.restype = i32
    L.trid_ragged_solve_f64.argtypes = [P, P, P, P, P, P, i32, i32, P]
    L.trid_ragged_solve_f64.restype = i32
    L.trid_ragged_solve_tol_f64.argtypes = [P, P, P, P, P, P, i32, ctypes.c_double, i32, P]
    L.trid_ragged_solve_tol_f64.restype = i32
    L.trid_server_create.argtypes = [ctypes.POINTER(P), ctypes.POINTER(ServerOpts)]
    L.trid_server_destroy.argtypes = [P]
    L.trid_server_register_operator.argtypes = [P, P, i64, P, P, P, ctypes.POINTER(ctypes.c_int32)]
    L.trid_server_release_operator.argtypes = [P, ctypes.c_int32]
    L.trid_server_submit.argtypes = [P, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]
    L.trid_server_wait.argtypes = [P, ctypes.c_uint64, i64, ctypes.POINTER(RequestResult)]
    L.trid_server_cancel.argtypes = [P, ctypes.c_uint64]
    L.trid_server_get_stats.argtypes = [P, ctypes.POINTER(ServerStats)]
    L.trid_server_create_adi.argtypes = [P, i64, P, ctypes.POINTER(ctypes.c_int32)]
    L.trid_server_read_adi.argtypes = [P, ctypes.c_int32, P]
    L.trid_server_release_adi.argtypes = [P, ctypes.c_int32]
    L.trid_server_host_alloc.argtypes = [P, ctypes.c_size_t, ctypes.POINTER(P)]
    L.trid_server_host_alloc.restype = i32
    L.trid_server_host_free.argtypes = [P, P]
    L.trid_server_host_free.restype = i32
    for name in ("trid_server_create_adi", "trid_server_read_adi", "trid_server_release_adi"):
        getattr(L, name).restype = i32
    for name in ("trid_server_create", "trid_server_destroy", "trid_server_register_operator",
                 "trid_server_release_operator", "trid_server_submit", "trid_server_wait",
                 "trid_server_cancel", "trid_server_get_stats"):
        getattr(L, name).restype = i32
    L.trid_adi_create.argtypes = [ctypes.POINTER(P), i64, P]
    L.trid_adi_create.restype = i32
    L.trid_adi_destroy.argtypes = [P]
    L.trid_adi_destroy.restype = i32
    L.trid_adi_grid.argtypes = [P]
    L.trid_adi_grid.restype = P
    L.trid_adi_step.argtypes = [P, ctypes.c_double, i32, P]
    L.trid_adi_step.restype = i32
    L.trid_plan_query.argtypes = [i32, i64, i64, i32, ctypes.POINTER(Opts),
                                  ctypes.POINTER(PlanInfo)]
    L.trid_plan_query.restype = i32
    _lib = L
    return L


def string(status: int) -> str:
    return lib().trid_status_string(int(status)).decode()


def check(status: int, where: str = "") -> None:
    if status != OK:
        raise TridError(status, where)
