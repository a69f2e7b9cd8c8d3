"""trid_solve_noalloc: the solve entry point that allocates nothing (its
workspace is the library's fixed device array). Called through the C ABI and
compared with trid_solve_f32 / trid_solve_f64 on a handle (same kernels, so
bit for bit) and with the FP64 Thomas oracle, on two calls in a row with the
second smaller, for each plan that needs a workspace: CR-PCR by separators
past the register sizes (FP32 and FP64) and the two-pass pcrGauss form.
Requests beyond the array's size are refused before anything is enqueued."""
import ctypes

import numpy as np
import pytest

from test_gpu_parity import check

ENOMEM = 2
SCRATCH = 64 << 20


def _noalloc(trid, method, dev, x, n, N, opts=None):
    import torch
    elem = 8 if x.dtype == torch.float64 else 4
    o = opts if opts is not None else trid.make_opts()
    ptrs = [ctypes.c_void_p(v.data_ptr()) for v in (*dev, x)]
    return trid.lib().trid_solve_noalloc(trid.METHODS[method], elem, *ptrs, n, N,
                                         ctypes.byref(o), None)


def test_noalloc_rejects_requests_beyond_the_scratch_array(trid):
    """Size check first: no device is needed to see the refusal."""
    # two-pass pcrGauss needs 3 n N elements: 65,536 x 128 FP32 = 96 MiB
    o = trid.make_opts(gaussian=16, flags=1)
    p = ctypes.c_void_p(256)
    assert 3 * 65536 * 128 * 4 > SCRATCH
    st = trid.lib().trid_solve_noalloc(trid.METHODS["pcrgauss"], 4, p, p, p, p, p, 65536, 128,
                                       ctypes.byref(o), None)
    assert st == ENOMEM
    assert trid.lib().trid_last_status() == ENOMEM
    # bad element size, null pointer
    assert trid.lib().trid_solve_noalloc(2, 2, p, p, p, p, p, 64, 1, None, None) == 1
    assert trid.lib().trid_solve_noalloc(2, 4, p, p, p, p, None, 64, 1, None, None) == 1


CASES = [
    # (method, dtype, [(n, N, opts kwargs), (smaller n, N, ...)])
    ("crpcr", np.float32, [(32768, 40, {}), (20000, 6, {})]),      # separators (FP32 > 16384)
    ("cr", np.float64, [(16384, 12, {}), (9000, 3, {})]),           # separators (FP64 > 8192)
    ("pcrgauss", np.float32, [(4096, 24, dict(gaussian=16, flags=1)),
                              (1000, 5, dict(gaussian=8, flags=1))]),
]


@pytest.mark.gpu
@pytest.mark.parametrize("method,dtype,calls", CASES, ids=[c[0] + "-" + c[1].__name__
                                                           for c in CASES])
def test_noalloc_equals_handle_solve(oracle, trid, method, dtype, calls):
    import torch
    h = trid.Handle()
    for k, (n, N, kw) in enumerate(calls):
        print(f"noalloc {method} {dtype.__name__} call {k}: n={n} N={N} {kw}", flush=True)
        a, b, c, d = oracle.generate("dominant", 70 + k, n, N, dtype)
        dev = [torch.from_numpy(v).cuda() for v in (a, b, c, d)]
        x_ref = torch.full_like(dev[0], float("nan"))
        x_na = torch.full_like(dev[0], float("nan"))
        o = trid.make_opts(**kw)
        h.solve(method, *dev, x_ref, n, N, o)
        st = _noalloc(trid, method, dev, x_na, n, N, o)
        assert st == 0, trid.lib().trid_status_string(st)
        torch.cuda.synchronize()
        xr, xn = x_ref.cpu().numpy(), x_na.cpu().numpy()
        assert np.array_equal(xr, xn), f"call {k}: differs from trid_solve"
        check(oracle, xn, a, b, c, d, n, dtype)
    h.close()
