// trid_adi.cpp -- resident ADI grids: Peaceman-Rachford steps of the 2D heat
// equation (kern_adi.hip) on a device grid that stays resident across steps.
// No counterpart in the reference; this is the caller the batched solve was
// built for (n row systems, n strided column systems per half step). Two
// scratch grids (the transposed right-hand side and the transposed solution
// of the column sweep) and the table of the operator's LU factors.
#include <hip/hip_runtime.h>

#include <new>

#include "trid.h"
#include "trid_launch.hpp"

struct trid_adi_s {
    int64_t n = 0;
    double* u = nullptr;
    double* t = nullptr;
    double* t2 = nullptr;
    double* fac = nullptr;
    bool owns_u = false;
};

extern "C" {

int trid_adi_create(trid_adi* out, int64_t n, double* u_device)
{
    if (!out || n < 2 || n > 16384) return TRID_EINVAL;
    *out = nullptr;
    trid_adi_s* a = new (std::nothrow) trid_adi_s();
    if (!a) return TRID_ENOMEM;
    a->n = n;
    const size_t grid = (size_t)n * (size_t)n * sizeof(double);
    hipError_t e = hipSuccess;
    if (u_device) {
        a->u = u_device;
    } else {
        e = hipMalloc(reinterpret_cast<void**>(&a->u), grid);
        a->owns_u = e == hipSuccess;
    }
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&a->t), grid);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&a->t2), grid);
    if (e == hipSuccess)
        e = hipMalloc(reinterpret_cast<void**>(&a->fac),
                      (size_t)trid::adi_npad(n) * sizeof(double));
    if (e != hipSuccess) {
        trid_adi_destroy(a);
        return TRID_ENOMEM;
    }
    *out = a;
    return TRID_OK;
}

int trid_adi_destroy(trid_adi a)
{
    if (!a) return TRID_EINVAL;
    if (a->owns_u && a->u) (void)hipFree(a->u);
    if (a->t) (void)hipFree(a->t);
    if (a->t2) (void)hipFree(a->t2);
    if (a->fac) (void)hipFree(a->fac);
    delete a;
    return TRID_OK;
}

double* trid_adi_grid(trid_adi a) { return a ? a->u : nullptr; }

int trid_adi_step(trid_adi a, double r, int steps, void* stream)
{
    if (!a || !(r > 0.0) || r > 1.0e6 || steps < 0) return TRID_EINVAL;
    trid::AdiArgs g{};
    g.u = a->u;
    g.t = a->t;
    g.t2 = a->t2;
    g.fac = a->fac;
    g.n = a->n;
    g.r = r;
    g.stream = static_cast<hipStream_t>(stream);
    return trid::launch_adi_steps(g, steps) == hipSuccess ? TRID_OK : TRID_EHIP;
}

}  // extern "C"
