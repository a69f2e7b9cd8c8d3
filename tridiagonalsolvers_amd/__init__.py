"""tridiagonalsolvers_amd -- batched tridiagonal solvers for AMD MI355X (gfx950).

Drop-in for the solve entry points of zhangyaobit/tridiagonalsolvers
(cr, pcr, crpcr, crNBC, crpcrNBC, crNBC32, pcrGaussWrapper) over the C ABI
of the in-tree ``libtrid.so`` (hand-written HIP kernels). See
``include/trid.h`` for the ABI.
"""
from .solvers import (  # noqa: F401
    METHODS,
    Handle,
    TridError,
    cr,
    crNBC,
    crNBC32,
    crpcr,
    crpcrNBC,
    device_count,
    generate,
    gtsv,
    make_opts,
    pcr,
    pcrGaussWrapper,
    plan_query,
    tridiagonalMKL,
)
from ._lib import LIB_PATH, lib  # noqa: F401
