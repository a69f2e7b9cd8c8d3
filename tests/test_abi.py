"""CPU tests of the C ABI: the library loads, exports every symbol that
include/*.h declares, and validates arguments before touching a device
(no solve is run here)."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        # only the extern "C" header declares C symbols; compat headers hold
        # inline C++ wrappers
        if 'extern "C"' not in text:
            continue
        for m in re.finditer(r"\b(trid_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(trid):
    names = declared_symbols()
    assert len(names) >= 24
    lib = trid.lib()
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_status_strings(trid):
    lib = trid.lib()
    for s in range(6):
        assert lib.trid_status_string(s).decode().startswith("TRID_")
    assert b"unknown" in lib.trid_status_string(99)


def test_host_dropins_reject_bad_arguments_without_a_device(trid):
    lib = trid.lib()
    a = np.ones(16, dtype=np.float32)
    p = ctypes.c_void_p(a.ctypes.data)
    null = ctypes.c_void_p(0)
    EINVAL = 1
    assert lib.trid_cr_f32(null, p, p, p, p, 16, 1) == EINVAL
    assert lib.trid_last_status() == EINVAL
    assert lib.trid_cr_f32(p, p, p, p, p, 1, 1) == EINVAL        # n < 2
    assert lib.trid_cr_f32(p, p, p, p, p, 16, -1) == EINVAL      # N < 0
    assert lib.trid_crpcr_f32(p, p, p, p, p, 16, 1, 3) == EINVAL   # rest not 2^k
    assert lib.trid_crpcr_f32(p, p, p, p, p, 16, 1, 16) == EINVAL  # rest > n/2
    assert lib.trid_crpcr_f32(p, p, p, p, p, 16, 1, 0) == EINVAL
    assert lib.trid_pcrgauss_f32(p, p, p, p, p, 16, 1, 3) == EINVAL
    assert lib.trid_pcrgauss_f32(p, p, p, p, p, 16, 1, 16) == EINVAL
    assert lib.trid_crpcrnbc_f64(p, p, p, p, p, 16, 1, 5) == EINVAL
    assert lib.trid_gtsv_f32(p, p, p, p, null, 16, 1) == EINVAL
    assert lib.trid_gtsv_f64(p, p, p, p, p, 0, 1) == EINVAL       # n < 1
    assert lib.trid_gtsv_f32(p, p, p, p, p, 16, -1) == EINVAL
    assert lib.trid_gtsv_f32(p, p, p, p, p, 1, 0) == 0            # empty batch


def test_device_api_rejects_bad_arguments_without_a_device(trid):
    h = trid.Handle()
    with pytest.raises(trid.TridError) as e:
        h.solve("crpcr", 0, 0, 0, 0, 0, 1024, 1)
    assert e.value.status == 1
    with pytest.raises(trid.TridError):
        h.solve("crpcr", 16, 16, 16, 16, 16, 1, 1)
    with pytest.raises(trid.TridError):
        h.solve("crpcr", 16, 16, 16, 16, 16, 64, 1, opts=trid.make_opts(flags=1))
    with pytest.raises(trid.TridError):
        h.solve(42, 16, 16, 16, 16, 16, 64, 1)
    # interleaved batches: stride below the system count, one row, x on a/b/c
    with pytest.raises(trid.TridError):
        h.solve_interleaved(16, 32, 48, 64, 80, 8, 10, 9)
    with pytest.raises(trid.TridError):
        h.solve_interleaved(16, 32, 48, 64, 80, 1, 10, 10)
    with pytest.raises(trid.TridError):
        h.solve_interleaved(16, 32, 48, 64, 48, 8, 10, 10)
    with pytest.raises(trid.TridError):
        h.solve_interleaved(0, 32, 48, 64, 80, 8, 10, 10)
    # pivoting solves: zero rows, x on a/b/c, interleaved stride below the
    # system count or zero, a null input; an empty batch is a no-op
    with pytest.raises(trid.TridError):
        h.solve_pivoting(16, 32, 48, 64, 80, 0, 10)
    with pytest.raises(trid.TridError):
        h.solve_pivoting(16, 32, 48, 64, 32, 8, 10)
    with pytest.raises(trid.TridError):
        h.solve_pivoting(16, 32, 48, 64, 80, 8, 10, batchStride=9)
    with pytest.raises(trid.TridError):
        h.solve_pivoting(16, 32, 48, 64, 80, 8, 10, batchStride=0)
    with pytest.raises(trid.TridError):
        h.solve_pivoting(0, 32, 48, 64, 80, 8, 10)
    # strided batches: stride below the system size, a method outside the
    # register family, flags; an empty batch is a no-op
    with pytest.raises(trid.TridError) as e:
        h.solve_strided("crpcr", 16, 32, 48, 64, 80, 64, 4, 63)
    assert e.value.status == 1
    with pytest.raises(trid.TridError) as e:
        h.solve_strided("crnbc", 16, 32, 48, 64, 80, 64, 4, 64)
    assert e.value.status == 4
    with pytest.raises(trid.TridError):
        h.solve_strided("crpcr", 16, 32, 48, 64, 80, 64, 4, 80, trid.make_opts(flags=1))
    h.solve_strided("crpcr", 16, 32, 48, 64, 80, 64, 0, 80)
    h.solve_pivoting(16, 32, 48, 64, 80, 8, 0)
    h.solve_pivoting(16, 32, 48, 64, 64, 1, 0, dtype="f64")
    h.close()


def test_plan_query_register_family(trid):
    q = trid.plan_query("crpcr", 1024, 65536, 4)
    assert q["family"] == 0 and q["per_lane"] * q["lanes"] == 1024
    assert q["per_lane"] == 16 and q["lanes"] == 64 and q["rest"] == 64
    q = trid.plan_query("cr", 512, 1, 4)
    assert q["rest"] == 2 and q["per_lane"] * q["lanes"] == 512
    q = trid.plan_query("pcr", 256, 1, 8)
    assert q["rest"] == 256
    q = trid.plan_query("crpcr", 1000, 3, 4)   # padded to 1024
    assert q["padded_size"] == 1024
    q = trid.plan_query("crpcr", 2, 1, 4)
    assert q["lanes"] == 2 and q["per_lane"] == 1 and q["rest"] == 2
    q = trid.plan_query("crpcr", 1024, 1, 4, trid.make_opts(rest=256))
    assert q["rest"] == 256
    q = trid.plan_query("crpcr", 64, 1, 4, trid.make_opts(lanes=8, per_lane=8))
    assert q["lanes"] == 8 and q["per_lane"] == 8
    with pytest.raises(trid.TridError):
        trid.plan_query("crpcr", 64, 1, 4, trid.make_opts(lanes=8, per_lane=4))
    with pytest.raises(trid.TridError):
        trid.plan_query("crpcr", 64, 1, 4, trid.make_opts(rest=64))
    with pytest.raises(trid.TridError):
        trid.plan_query("crpcr", 64, 1, 3)


def test_plan_query_separator_family(trid):
    # cr / crpcr past the register sizes: chunks of 1,024 rows (16 x 64 lanes)
    for n, es in ((16385, 4), (1 << 20, 4), (1 << 30, 4), (8193, 8), (1 << 20, 8)):
        for m in ("cr", "crpcr", "auto"):
            q = trid.plan_query(m, n, 3, es)
            assert q["family"] == 3 and q["per_lane"] == 16 and q["lanes"] == 64, (m, n, es)
            assert q["rest"] == (2 if m == "cr" else 64)
    assert trid.plan_query("crpcr", 16384, 3, 4)["family"] == 0
    assert trid.plan_query("crpcr", 8192, 3, 8)["family"] == 0
    # multi-wave systems: PCR on one equation per lane of the workgroup
    q = trid.plan_query("crpcr", 8192, 4096, 8)
    assert q["lanes"] == 512 and q["per_lane"] == 16 and q["rest"] == 512
    q = trid.plan_query("crpcr", 8192, 4096, 4)
    assert q["lanes"] * q["per_lane"] == 8192 and q["rest"] == q["lanes"]
    q = trid.plan_query("crpcr", 1 << 20, 1, 4, trid.make_opts(rest=512))
    assert q["family"] == 3 and q["rest"] == 512
    for bad, status in ((trid.make_opts(rest=1024), 4), (trid.make_opts(rest=96), 1),
                        (trid.make_opts(lanes=64, per_lane=16), 1)):
        with pytest.raises(trid.TridError) as e:
            trid.plan_query("crpcr", 1 << 20, 1, 4, bad)
        assert e.value.status == status
    with pytest.raises(trid.TridError) as e:     # PCR has no separator form
        trid.plan_query("pcr", 1 << 20, 1, 4)
    assert e.value.status == 4


def test_plan_query_nbc_and_gauss(trid):
    q = trid.plan_query("crnbc", 1024, 1, 4)
    assert q["family"] == 1 and q["lds_pad"] > 0
    q = trid.plan_query("crnbc32", 1024, 1, 4)
    assert q["family"] == 1 and q["lds_pad"] == 0
    q = trid.plan_query("crpcrnbc", 1024, 1, 4, trid.make_opts(rest=128))
    assert q["family"] == 1 and q["rest"] == 128
    q = trid.plan_query("pcrgauss", 1024, 1, 4, trid.make_opts(gaussian=64))
    assert q["family"] == 2 and q["gaussian"] == 64


def test_ragged_plan_rejects_bad_sizes_without_a_device(trid):
    import tridiagonalsolvers_amd.ragged as R
    with pytest.raises(trid.TridError) as e:
        R.RaggedPlan(np.array([0, 5, 6, 10], dtype=np.int64), table=[])   # a size of 1
    assert e.value.status == 1
    with pytest.raises(trid.TridError) as e:
        R.RaggedPlan(np.array([0, 5, 3], dtype=np.int64), table=[])       # negative size
    assert e.value.status == 1
    with pytest.raises(trid.TridError) as e:
        R.RaggedPlan(np.array([0, 20000], dtype=np.int64), table=[])      # above 16384
    assert e.value.status == 4
    bad = [dict(size=64, per_lane=16, lanes=4, rest=4)]                    # per_lane > 8
    with pytest.raises(trid.TridError):
        R.RaggedPlan(np.array([0, 64], dtype=np.int64), table=bad)
    bad = [dict(size=64, per_lane=8, lanes=8, rest=16)]                    # rest > lanes
    with pytest.raises(trid.TridError):
        R.RaggedPlan(np.array([0, 64], dtype=np.int64), table=bad)
    # 8192-unknown systems run at 16 unknowns per lane (8 waves of 64 lanes);
    # below, 16 per lane needs at least two waves
    for bad in ([dict(size=8192, per_lane=8, lanes=1024, rest=512)],
                [dict(size=8192, per_lane=16, lanes=256, rest=256)],
                [dict(size=1024, per_lane=16, lanes=64, rest=64)]):
        n = bad[0]["size"]
        with pytest.raises(trid.TridError) as e:
            R.RaggedPlan(np.array([0, n], dtype=np.int64), table=bad)
        assert e.value.status == 1


def test_ragged_tolerance_arguments_without_a_device(trid):
    lib = trid.lib()
    null = ctypes.c_void_p(0)
    for tol, flags in ((-1.0, 0), (float("nan"), 0), (float("inf"), 0), (1e-12, 2)):
        assert lib.trid_ragged_solve_tol_f64(null, null, null, null, null, null, 2,
                                             ctypes.c_double(tol), flags, null) == 1


def test_autotune_candidates_are_valid_configs():
    from tridiagonalsolvers_amd.autotune import candidates
    for k in range(1, 15):
        size = 1 << k
        cs = candidates(size)
        assert cs, size
        for c in cs:
            assert c["per_lane"] * c["lanes"] == size
            assert c["per_lane"] <= (16 if size >= 2048 else 8)
            assert 2 <= c["rest"] <= c["lanes"] <= 1024
            if size <= 512:
                assert c["lanes"] <= 64
            else:
                assert c["lanes"] >= 128
                assert c["per_lane"] == 16 if size >= 8192 else c["per_lane"] in (8, 16)


def test_log_uniform_sizes_are_seeded_and_in_range():
    import tridiagonalsolvers_amd.ragged as R
    s1 = R.log_uniform_sizes(10000, 16, 8192, seed=1)
    s2 = R.log_uniform_sizes(10000, 16, 8192, seed=1)
    assert np.array_equal(s1, s2)
    assert s1.min() >= 16 and s1.max() <= 8192
    # log-uniform: about a third of the systems at or above 1024
    frac = float(np.mean(s1 >= 1024))
    assert 0.28 < frac < 0.39


def test_adi_grid_sizes_are_checked_without_a_device(trid):
    # grids of 2 .. 16384 points a side; others are refused before anything
    # is allocated
    import ctypes
    from tridiagonalsolvers_amd._lib import lib
    h = ctypes.c_void_p()
    assert lib().trid_adi_create(ctypes.byref(h), 16385, None) == 1      # TRID_EINVAL
    assert lib().trid_adi_create(ctypes.byref(h), 1, None) == 1


def test_device_tensors_are_checked_before_the_c_abi(trid):
    """Torch tensors handed to the device API are checked on the host: a
    tensor in host memory, a short one, a non-contiguous one or mixed
    dtypes raise before any kernel could see a host pointer or read past
    an array's end."""
    import torch
    import tridiagonalsolvers_amd.ragged as R
    h = trid.Handle()
    cpu = [torch.zeros(64, dtype=torch.float32) for _ in range(5)]
    with pytest.raises(ValueError, match="not on a GPU"):
        h.solve("crpcr", *cpu, 64, 1)
    with pytest.raises(ValueError, match="not on a GPU"):
        h.solve_interleaved(*cpu, 64, 1)
    with pytest.raises(ValueError, match="not on a GPU"):
        trid.generate(0, 1, *cpu[:4], 64, 1)
    with pytest.raises(TypeError):
        h.solve("crpcr", *cpu[:4], torch.zeros(64, dtype=torch.float64), 64, 1)
    with pytest.raises(TypeError):
        h.solve("crpcr", *(torch.zeros(64, dtype=torch.float16) for _ in range(5)), 64, 1)
    with pytest.raises(ValueError, match="not on a GPU"):
        h.solve_refined("crpcr", *(torch.zeros(64, dtype=torch.float64) for _ in range(5)),
                        64, 1)
    with pytest.raises(TypeError):
        h.solve_refined("crpcr", *cpu, 64, 1)             # FP64 only
    # the ragged entry points use the same checks (a plan itself needs a device)
    with pytest.raises(ValueError, match="not on a GPU"):
        R.residual(*(torch.zeros(56, dtype=torch.float64) for _ in range(5)),
                   torch.tensor([0, 16, 56]), 2, torch.zeros(2, dtype=torch.float64))
    # the checks that need no device: lengths and layout, with a stand-in
    # object that reports itself as a GPU tensor

    class FakeGpu:
        def __init__(self, n, dtype=torch.float32, contiguous=True):
            self.dtype, self._n, self._c, self.is_cuda = dtype, n, contiguous, True

        def numel(self):
            return self._n

        def is_contiguous(self):
            return self._c

        def data_ptr(self):
            return 256

    with pytest.raises(ValueError, match="63 elements, 64 needed"):
        h.solve("crpcr", *(FakeGpu(64) for _ in range(4)), FakeGpu(63), 64, 1)
    with pytest.raises(ValueError, match="not contiguous"):
        h.solve("crpcr", *(FakeGpu(64) for _ in range(4)), FakeGpu(64, contiguous=False), 64, 1)
    # strided: (N - 1) * stride + n elements; interleaved: (n - 1) * stride + N
    with pytest.raises(ValueError, match="needed"):
        h.solve_strided("crpcr", *(FakeGpu(3 * 80 + 63) for _ in range(5)), 64, 4, 80)
    with pytest.raises(ValueError, match="needed"):
        h.solve_interleaved(*(FakeGpu(63 * 10 + 7) for _ in range(5)), 64, 8, 10)
    with pytest.raises(ValueError, match="info"):
        h.solve_pivoting(*(FakeGpu(640) for _ in range(5)), 64, 10,
                         info=FakeGpu(9, dtype=torch.int32))
    with pytest.raises(TypeError, match="info"):
        h.solve_pivoting(*(FakeGpu(640) for _ in range(5)), 64, 10,
                         info=FakeGpu(10, dtype=torch.int64))
    h.close()


def test_bucket_table_loader_reads_the_shipped_tables(trid, tmp_path):
    """trid_bucket_table_load (the C reader of the autotune JSON) gives the
    same entries as the Python reader for both shipped tables, and refuses
    malformed files, missing keys and invalid configurations."""
    import json

    import tridiagonalsolvers_amd.ragged as R
    from tridiagonalsolvers_amd._lib import BucketCfg
    lib = trid.lib()

    def load(path, cap=32):
        arr = (BucketCfg * max(cap, 1))()
        cnt = ctypes.c_int(-1)
        st = lib.trid_bucket_table_load(str(path).encode(), ctypes.cast(arr, ctypes.c_void_p),
                                        cap, ctypes.byref(cnt))
        return st, cnt.value, [dict(size=arr[i].size, per_lane=arr[i].per_lane,
                                    lanes=arr[i].lanes, rest=arr[i].rest)
                               for i in range(min(cnt.value, cap))]

    for path in (R.DEFAULT_TABLE, R.DEFAULT_TABLE_FP64):
        st, cnt, got = load(path)
        assert st == 0 and cnt >= 10
        assert got == R.load_table(path)
        st, cnt2, part = load(path, cap=3)          # count is the file's, 3 stored
        assert st == 0 and cnt2 == cnt and part == got[:3]
    good = {"device": "x \"quoted\" \\u00e9", "nested": {"a": [1, 2.5e3, True, None]},
            "buckets": [{"size": 64, "per_lane": 8, "lanes": 8, "rest": 8, "ms": 0.1}]}
    p = tmp_path / "good.json"
    p.write_text(json.dumps(good))
    assert load(p)[:2] == (0, 1)
    bad = [
        "",                                                         # empty
        "[1, 2]",                                                   # not an object
        '{"buckets": [{"size": 64, "per_lane": 8, "lanes": 8}]}',   # rest missing
        '{"buckets": [{"size": 64, "per_lane": 8, "lanes": 4, "rest": 4}]}',  # 8*4 != 64
        '{"buckets": [{"size": 64.5, "per_lane": 8, "lanes": 8, "rest": 8}]}',
        '{"buckets": [{"size": 64, "per_lane": 8, "lanes": 8, "rest": 8}]',   # unterminated
        '{"other": 1}',                                             # no buckets
        '{"buckets": [{"size": 64, "per_lane": 8, "lanes": 8, "rest": 8}]} x',  # trailing text
    ]
    for k, text in enumerate(bad):
        q = tmp_path / f"bad{k}.json"
        q.write_text(text)
        assert load(q)[0] == 1, text
    assert load(tmp_path / "missing.json")[0] == 1
