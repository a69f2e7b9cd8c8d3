/*
 * trid_compat.h -- the reference's solver names as inline C++ overloads.
 *
 * zhangyaobit/tridiagonalsolvers declares its whole solver API as forward
 * declarations (test_tridiagonal.cpp:17-26): free C++ functions over host
 * pointers with a compile-time element type T and a void return. Including
 * this header gives code written against those declarations the same names,
 * argument order and meaning, for both float and double, forwarding to the
 * C ABI of libtrid.so (trid.h). Like the reference they return nothing; the
 * status of the call is available from trid_last_status(). Unlike the
 * reference nothing exits the process, prints, or runs a timing loop.
 */
#ifndef TRID_COMPAT_H
#define TRID_COMPAT_H

#include "trid.h"

#define TRID_COMPAT_7(NAME, CFN)                                                      \
    inline void NAME(float* a, float* b, float* c, float* d, float* x, int systemSize, \
                     int numSystems)                                                  \
    {                                                                                 \
        (void)CFN##_f32(a, b, c, d, x, systemSize, numSystems);                       \
    }                                                                                 \
    inline void NAME(double* a, double* b, double* c, double* d, double* x,           \
                     int systemSize, int numSystems)                                  \
    {                                                                                 \
        (void)CFN##_f64(a, b, c, d, x, systemSize, numSystems);                       \
    }

#define TRID_COMPAT_8(NAME, CFN, EXTRA)                                                \
    inline void NAME(float* a, float* b, float* c, float* d, float* x, int systemSize, \
                     int numSystems, int EXTRA)                                       \
    {                                                                                 \
        (void)CFN##_f32(a, b, c, d, x, systemSize, numSystems, EXTRA);                \
    }                                                                                 \
    inline void NAME(double* a, double* b, double* c, double* d, double* x,           \
                     int systemSize, int numSystems, int EXTRA)                       \
    {                                                                                 \
        (void)CFN##_f64(a, b, c, d, x, systemSize, numSystems, EXTRA);                \
    }

TRID_COMPAT_7(cr, trid_cr)                                   /* cr_app.cu:4          */
TRID_COMPAT_7(pcr, trid_pcr)                                 /* pcr_app.cu:5         */
TRID_COMPAT_8(crpcr, trid_crpcr, restSystemSize)             /* crpcr_app.cu:5       */
TRID_COMPAT_7(crNBC, trid_crnbc)                             /* crNBC.cu:123         */
TRID_COMPAT_8(crpcrNBC, trid_crpcrnbc, restSystemSize)       /* crNBC.cu:344         */
TRID_COMPAT_7(crNBC32, trid_crnbc32)                         /* crNBC32banks.cu:123  */
TRID_COMPAT_8(pcrGaussWrapper, trid_pcrgauss, gaussian)      /* pcrGauss_app.cu:164  */

/* tridiagonalMKL.cpp:5: ?gtsv per system, solving in place in d and then
 * swapping the d and x pointers, so that afterwards x points at the
 * solution. Kept here: the solution is written into the d buffer and the
 * pointers are swapped. a, b, c are left unmodified (?gtsv overwrites them). */
inline void tridiagonalMKL(float* a, float* b, float* c, float*& d, float*& x, int system_size,
                           int num_systems)
{
    if (trid_gtsv_f32(a, b, c, d, d, system_size, num_systems) != TRID_OK) return;
    float* t = x;
    x = d;
    d = t;
}
inline void tridiagonalMKL(double* a, double* b, double* c, double*& d, double*& x,
                           int system_size, int num_systems)
{
    if (trid_gtsv_f64(a, b, c, d, d, system_size, num_systems) != TRID_OK) return;
    double* t = x;
    x = d;
    d = t;
}

/* tridiagonalMKL.cpp:28: the same per-system ?gtsv as tridiagonalMKL, spread
 * by the reference over 4 OpenMP threads that each take num_systems / 4
 * systems (the last num_systems mod 4 are left unsolved). Here every system
 * is solved, by the same GPU pivoting solve as tridiagonalMKL, with the same
 * in-d result and d/x pointer swap. */
inline void multiThreadTridiagonalMKL(float* a, float* b, float* c, float*& d, float*& x,
                                      int system_size, int num_systems)
{
    tridiagonalMKL(a, b, c, d, x, system_size, num_systems);
}
inline void multiThreadTridiagonalMKL(double* a, double* b, double* c, double*& d, double*& x,
                                      int system_size, int num_systems)
{
    tridiagonalMKL(a, b, c, d, x, system_size, num_systems);
}

#undef TRID_COMPAT_7
#undef TRID_COMPAT_8

#endif /* TRID_COMPAT_H */
