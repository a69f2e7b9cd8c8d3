/*
 * trid_oracle.c -- CPU oracle for the batched tridiagonal solve.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: it is linked or loaded
 * by tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of
 * bench.py, and by nothing else. The shipped library never calls it.
 *
 * What it restates (citations are relative to the reference repository
 * zhangyaobit/tridiagonalsolvers, directory tridiagonalsolvers/):
 *
 *  - oracle_thomas_*: the serial Thomas algorithm of tridiagonal_gold.h:26-44
 *    (`serial`), with the same order of operations:
 *        c'[0] = c[0]/b[0];  d'[0] = d[0]/b[0]
 *        c'[i] = c[i] / (b[i] - a[i]*c'[i-1])
 *        d'[i] = (d[i] - d'[i-1]*a[i]) / (b[i] - a[i]*c'[i-1])
 *        x[n-1] = d'[n-1];  x[i] = d'[i] - c'[i]*x[i+1]
 *    a[0] is never read and c[n-1] is taken as 0 (tridiagonal_gold.h:28).
 *    Unlike the gold it is non-destructive (the gold overwrites c and d), and
 *    the per-system loop (tridiagonal_gold.h:47-53) runs over ALL systems
 *    with OpenMP (the gold's 4-thread variant, tridiagonal_gold.h:56-72,
 *    silently drops numSystems mod 4 systems; that is not reproduced).
 *    Systems may have different sizes (offsets array) so the same code is the
 *    oracle for ragged batches.
 *
 *  - oracle_compare_*: the per-system RMS of tridiagonal_gold.h:97-119 with
 *    the NaN-is-failure rule of tridiagonal_gold.h:131-148, plus the primary
 *    metric (max over systems of the relative infinity-norm error) and the
 *    relative residual ||d - T x||_inf / ||d||_inf in long double.
 *
 *  - oracle_gtsv_*: partial pivoting (LAPACK ?gtsv), the algorithm of the
 *    reference's MKL baseline (tridiagonalMKL.cpp:5-26), in FP32 and FP64.
 *
 *  - oracle_generate_*: a seeded, counter-based input generator. The
 *    reference generator (tridiagonal_gold.h:75-94) uses unseeded rand(), so
 *    it cannot be reproduced; this one is defined here and restated
 *    independently in numpy (oracle/oracle.py) and in the product's GPU
 *    generator. The "refgen" kind keeps the reference's value ranges
 *    (b in [8,9], a in [3,4], c in [2,3], d in [5,6]).
 *
 * Build: oracle/Makefile (gcc, OpenMP, -ffp-contract=off so that the
 * generator's double arithmetic is the plain IEEE sequence written here).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* ------------------------------------------------------------------------ */
/* Thomas                                                                    */
/* ------------------------------------------------------------------------ */

/* One system, FP64 arithmetic. cp/dp are caller scratch of length n. */
#define THOMAS_BODY(TIN, TW)                                                  \
    {                                                                         \
        if (n <= 0) return;                                                   \
        if (n == 1) { x[0] = (TW)d[0] / (TW)b[0]; return; }                   \
        cp[0] = (TW)c[0] / (TW)b[0];                                          \
        dp[0] = (TW)d[0] / (TW)b[0];                                          \
        for (int64_t i = 1; i < n; i++) {                                     \
            TW ci = (i == n - 1) ? (TW)0 : (TW)c[i];                          \
            TW ai = (TW)a[i];                                                 \
            cp[i] = ci / ((TW)b[i] - ai * cp[i - 1]);                         \
            dp[i] = ((TW)d[i] - dp[i - 1] * ai) / ((TW)b[i] - ai * cp[i - 1]);\
        }                                                                     \
        x[n - 1] = dp[n - 1];                                                 \
        for (int64_t i = n - 2; i >= 0; i--) x[i] = dp[i] - cp[i] * x[i + 1]; \
    }

static void thomas1_f64_f32in(const float* a, const float* b, const float* c,
                              const float* d, double* x, double* cp, double* dp,
                              int64_t n) THOMAS_BODY(float, double)

static void thomas1_f64_f64in(const double* a, const double* b, const double* c,
                              const double* d, double* x, double* cp, double* dp,
                              int64_t n) THOMAS_BODY(double, double)

static void thomas1_f32(const float* a, const float* b, const float* c,
                        const float* d, float* x, float* cp, float* dp,
                        int64_t n) THOMAS_BODY(float, float)

static int nthreads_or_max(int threads)
{
    return threads > 0 ? threads : omp_get_max_threads();
}

/* Uniform batch: system s occupies [s*n, (s+1)*n). */
#define UNIFORM_DRIVER(NAME, TIN, TOUT, TW, ONE)                              \
    int NAME(const TIN* a, const TIN* b, const TIN* c, const TIN* d, TOUT* x,  \
             int64_t n, int64_t num_systems, int threads)                     \
    {                                                                         \
        if (n <= 0 || num_systems < 0) return 1;                              \
        int nt = nthreads_or_max(threads);                                    \
        int fail = 0;                                                         \
        _Pragma("omp parallel num_threads(nt)")                               \
        {                                                                     \
            TW* cp = (TW*)malloc(sizeof(TW) * (size_t)n);                     \
            TW* dp = (TW*)malloc(sizeof(TW) * (size_t)n);                     \
            if (!cp || !dp) {                                                 \
                _Pragma("omp atomic write") fail = 1;                         \
            } else {                                                          \
                _Pragma("omp for schedule(static)")                           \
                for (int64_t s = 0; s < num_systems; s++) {                   \
                    int64_t o = s * n;                                        \
                    ONE(a + o, b + o, c + o, d + o, x + o, cp, dp, n);        \
                }                                                             \
            }                                                                 \
            free(cp);                                                         \
            free(dp);                                                         \
        }                                                                     \
        return fail;                                                          \
    }

UNIFORM_DRIVER(oracle_thomas_f64_from_f32, float, double, double, thomas1_f64_f32in)
UNIFORM_DRIVER(oracle_thomas_f64, double, double, double, thomas1_f64_f64in)
UNIFORM_DRIVER(oracle_thomas_f32, float, float, float, thomas1_f32)

/* Ragged batch: system s occupies [offsets[s], offsets[s+1]). */
#define RAGGED_DRIVER(NAME, TIN, TOUT, TW, ONE)                               \
    int NAME(const TIN* a, const TIN* b, const TIN* c, const TIN* d, TOUT* x,  \
             const int64_t* offsets, int64_t num_systems, int threads)        \
    {                                                                         \
        int64_t nmax = 0;                                                     \
        for (int64_t s = 0; s < num_systems; s++) {                           \
            int64_t len = offsets[s + 1] - offsets[s];                        \
            if (len <= 0) return 1;                                           \
            if (len > nmax) nmax = len;                                       \
        }                                                                     \
        int nt = nthreads_or_max(threads);                                    \
        int fail = 0;                                                         \
        _Pragma("omp parallel num_threads(nt)")                               \
        {                                                                     \
            TW* cp = (TW*)malloc(sizeof(TW) * (size_t)(nmax + 1));            \
            TW* dp = (TW*)malloc(sizeof(TW) * (size_t)(nmax + 1));            \
            if (!cp || !dp) {                                                 \
                _Pragma("omp atomic write") fail = 1;                         \
            } else {                                                          \
                _Pragma("omp for schedule(dynamic, 64)")                      \
                for (int64_t s = 0; s < num_systems; s++) {                   \
                    int64_t o = offsets[s], len = offsets[s + 1] - o;         \
                    ONE(a + o, b + o, c + o, d + o, x + o, cp, dp, len);      \
                }                                                             \
            }                                                                 \
            free(cp);                                                         \
            free(dp);                                                         \
        }                                                                     \
        return fail;                                                          \
    }

RAGGED_DRIVER(oracle_thomas_ragged_f64_from_f32, float, double, double, thomas1_f64_f32in)
RAGGED_DRIVER(oracle_thomas_ragged_f64, double, double, double, thomas1_f64_f64in)

int oracle_max_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------------ */
/* Partial pivoting (LAPACK ?gtsv, one right-hand side)                      */
/* ------------------------------------------------------------------------ */

/*
 * The reference's CPU baseline calls MKL ?gtsv on each system
 * (tridiagonalMKL.cpp:5-26; &a[s*n+1] is passed as the sub-diagonal, so
 * a[first] is skipped, and c[last] is never read). MKL is not in this image.
 * This is the published LAPACK algorithm (?gtsv, the NRHS = 1 branch)
 * restated in the element type itself (float = sgtsv, double = dgtsv) and
 * made non-destructive: D, U, W are caller scratch (diagonal, first and
 * second super-diagonal of U after the row interchanges), the right-hand side
 * is carried in x.
 *   for i = 0 .. n-2, with l = a[i+1] the entry under the pivot:
 *     |D_i| >= |l|: D_i == 0 -> info = i+1 (stop); else f = l/D_i,
 *                   D_{i+1} -= f U_i, x_{i+1} -= f x_i, W_i = 0
 *     else:         f = D_i/l, D_i = l, t = D_{i+1}, D_{i+1} = U_i - f t,
 *                   W_i = U_{i+1}, U_{i+1} = -f W_i (both skipped at i = n-2),
 *                   U_i = t, swap x_i, x_{i+1} with x_{i+1} = x_i - f x_{i+1}
 *   D_{n-1} == 0 -> info = n
 *   x_{n-1} /= D_{n-1}; x_{n-2} = (x_{n-2} - U x_{n-1}) / D;
 *   x_i = (x_i - U_i x_{i+1} - W_i x_{i+2}) / D_i
 * Returns info (0 = solved); with info != 0 x is set to NaN, the product's
 * rule for a singular system (LAPACK leaves it uncomputed).
 */
#define GTSV_ONE(NAME, TT, ABS, QNAN)                                         \
    static int64_t NAME(const TT* a, const TT* b, const TT* c, const TT* d,   \
                        TT* x, TT* D, TT* U, TT* W, int64_t n)                \
    {                                                                         \
        int64_t info = 0;                                                     \
        for (int64_t i = 0; i < n; i++) {                                     \
            D[i] = b[i];                                                      \
            U[i] = (i < n - 1) ? c[i] : (TT)0;                                \
            W[i] = (i < n - 1) ? a[i + 1] : (TT)0;                            \
            x[i] = d[i];                                                      \
        }                                                                     \
        for (int64_t i = 0; i < n - 1 && info == 0; i++) {                    \
            if (ABS(D[i]) >= ABS(W[i])) {                                     \
                if (D[i] != (TT)0) {                                          \
                    TT f = W[i] / D[i];                                       \
                    D[i + 1] = D[i + 1] - f * U[i];                           \
                    x[i + 1] = x[i + 1] - f * x[i];                           \
                } else {                                                      \
                    info = i + 1;                                             \
                }                                                             \
                W[i] = (TT)0;                                                 \
            } else {                                                          \
                TT f = D[i] / W[i];                                           \
                D[i] = W[i];                                                  \
                TT t = D[i + 1];                                              \
                D[i + 1] = U[i] - f * t;                                      \
                if (i < n - 2) {                                              \
                    W[i] = U[i + 1];                                          \
                    U[i + 1] = -f * W[i];                                     \
                } else {                                                      \
                    W[i] = (TT)0;                                             \
                }                                                             \
                U[i] = t;                                                     \
                t = x[i];                                                     \
                x[i] = x[i + 1];                                              \
                x[i + 1] = t - f * x[i + 1];                                  \
            }                                                                 \
        }                                                                     \
        if (info == 0 && D[n - 1] == (TT)0) info = n;                         \
        if (info != 0) {                                                      \
            for (int64_t i = 0; i < n; i++) x[i] = QNAN;                      \
            return info;                                                      \
        }                                                                     \
        x[n - 1] = x[n - 1] / D[n - 1];                                       \
        if (n > 1) x[n - 2] = (x[n - 2] - U[n - 2] * x[n - 1]) / D[n - 2];    \
        for (int64_t i = n - 3; i >= 0; i--)                                  \
            x[i] = (x[i] - U[i] * x[i + 1] - W[i] * x[i + 2]) / D[i];         \
        return 0;                                                             \
    }

GTSV_ONE(gtsv1_f32, float, fabsf, nanf(""))
GTSV_ONE(gtsv1_f64, double, fabs, nan(""))

/* Uniform batch, system s at [s*n, (s+1)*n); info may be NULL. */
#define GTSV_DRIVER(NAME, TT, ONE)                                            \
    int NAME(const TT* a, const TT* b, const TT* c, const TT* d, TT* x,      \
             int64_t n, int64_t num_systems, int32_t* info, int threads)      \
    {                                                                         \
        if (n <= 0 || num_systems < 0) return 1;                              \
        int nt = nthreads_or_max(threads);                                    \
        int fail = 0;                                                         \
        _Pragma("omp parallel num_threads(nt)")                               \
        {                                                                     \
            TT* D = (TT*)malloc(sizeof(TT) * (size_t)n * 3);                  \
            if (!D) {                                                         \
                _Pragma("omp atomic write") fail = 1;                         \
            } else {                                                          \
                _Pragma("omp for schedule(static)")                           \
                for (int64_t s = 0; s < num_systems; s++) {                   \
                    int64_t o = s * n;                                        \
                    int64_t r = ONE(a + o, b + o, c + o, d + o, x + o, D,     \
                                    D + n, D + 2 * n, n);                     \
                    if (info) info[s] = (int32_t)r;                           \
                }                                                             \
            }                                                                 \
            free(D);                                                          \
        }                                                                     \
        return fail;                                                          \
    }

GTSV_DRIVER(oracle_gtsv_f32, float, gtsv1_f32)
GTSV_DRIVER(oracle_gtsv_f64, double, gtsv1_f64)

/* ------------------------------------------------------------------------ */
/* Comparison and residual                                                   */
/* ------------------------------------------------------------------------ */

/*
 * Per-system metrics of a candidate x (as double) against the FP64 oracle
 * xref. For each system s:
 *   relerr[s] = max_i |x_i - xref_i| / max_i |xref_i|   (NaN if any x_i is NaN)
 *   rms[s]    = sqrt(mean_i (x_i - xref_i)^2)            (tridiagonal_gold.h:97-119)
 * Returns the number of systems that are NaN or have rms > rms_eps
 * (tridiagonal_gold.h:131-148), which is the reference's pass/fail count.
 */
int64_t oracle_compare_ragged(const double* x, const double* xref,
                              const int64_t* offsets, int64_t num_systems,
                              double rms_eps, double* relerr, double* rms)
{
    int64_t failures = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : failures)
    for (int64_t s = 0; s < num_systems; s++) {
        int64_t o = offsets[s], len = offsets[s + 1] - o;
        double maxdiff = 0.0, maxref = 0.0, sq = 0.0;
        int nan = 0;
        for (int64_t i = 0; i < len; i++) {
            double xi = x[o + i], ri = xref[o + i];
            double df = xi - ri;
            if (xi != xi || ri != ri) nan = 1;
            if (fabs(df) > maxdiff) maxdiff = fabs(df);
            if (fabs(ri) > maxref) maxref = fabs(ri);
            sq += df * df;
        }
        double r = sqrt(sq / (double)len);
        double e = maxref > 0.0 ? maxdiff / maxref : maxdiff;
        if (nan) { r = NAN; e = NAN; }
        if (relerr) relerr[s] = e;
        if (rms) rms[s] = r;
        if (nan || r > rms_eps) failures++;
    }
    return failures;
}

/*
 * Relative residual per system: ||d - T x||_inf / ||d||_inf, evaluated in
 * long double with a[first] and c[last] taken as zero. Inputs are double
 * (callers promote float data first).
 */
void oracle_residual_ragged(const double* a, const double* b, const double* c,
                            const double* d, const double* x,
                            const int64_t* offsets, int64_t num_systems,
                            double* relres)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t s = 0; s < num_systems; s++) {
        int64_t o = offsets[s], len = offsets[s + 1] - o;
        long double rmax = 0.0L, dmax = 0.0L;
        for (int64_t i = 0; i < len; i++) {
            long double r = (long double)d[o + i] - (long double)b[o + i] * x[o + i];
            if (i > 0) r -= (long double)a[o + i] * x[o + i - 1];
            if (i < len - 1) r -= (long double)c[o + i] * x[o + i + 1];
            if (fabsl(r) > rmax || r != r) rmax = (r != r) ? (long double)NAN : fabsl(r);
            if (fabsl((long double)d[o + i]) > dmax) dmax = fabsl((long double)d[o + i]);
        }
        relres[s] = (double)(dmax > 0.0L ? rmax / dmax : rmax);
    }
}

/* ------------------------------------------------------------------------ */
/* Seeded generator                                                          */
/* ------------------------------------------------------------------------ */

/* splitmix64 output function (Steele, Lea, Flood 2014; constants as
 * published by Vigna). */
static inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Stream key for (seed, stream id). Stream ids: 0=a 1=b 2=c 3=d 4=per-system
 * scalars (ADI r, ragged sizes). */
uint64_t oracle_stream_key(uint64_t seed, uint64_t stream)
{
    return mix64(seed ^ mix64(stream + 1));
}

/* Uniform double in [0,1) with 53 random bits, for element index g. */
static inline double u01(uint64_t key, uint64_t g)
{
    return (double)(mix64(key ^ g) >> 11) * 0x1.0p-53;
}

double oracle_u01(uint64_t seed, uint64_t stream, uint64_t g)
{
    return u01(oracle_stream_key(seed, stream), g);
}

enum { GEN_DOMINANT = 0, GEN_WEAK = 1, GEN_REFGEN = 2, GEN_PAIRED = 3, GEN_GENERAL = 4 };

/*
 * Fill one flat batch. Element g of the flat arrays belongs to the system
 * whose range [offsets[s], offsets[s+1]) contains it; the first element of
 * each system has a = 0 and the last has c = 0.
 *   DOMINANT: a,c ~ U[-1,1], b = |a|+|c| + U[1,2],      d ~ U[-1,1]
 *   WEAK:     a,c ~ U[-1,1], b = |a|+|c| + U[0.01,0.1], d ~ U[-1,1]
 *   REFGEN:   b = 8+U, a = 3+U, c = 2+U, d = 5+U (reference value ranges)
 *   PAIRED:   not diagonally dominant, well conditioned after row
 *             interchanges: within system index j, even rows a = 0.25(2U-1),
 *             c = 1+U; odd rows a = 1+U, c = 0.25(2U-1); b = 0.001(2U-1)
 *             (b = 1+U on the last row when it is even); d = 2U-1
 *   GENERAL:  a, b, c, d = 2U-1 (not dominant; may be ill-conditioned)
 * All arithmetic is double; each value is rounded once to the output type.
 * b uses the already-rounded a and c.
 */
#define GEN_DRIVER(NAME, TT)                                                  \
    int NAME(int kind, uint64_t seed, TT* a, TT* b, TT* c, TT* d,             \
             const int64_t* offsets, int64_t num_systems)                     \
    {                                                                         \
        if (kind < 0 || kind > 4) return 1;                                   \
        uint64_t ka = oracle_stream_key(seed, 0), kb = oracle_stream_key(seed, 1); \
        uint64_t kc = oracle_stream_key(seed, 2), kd = oracle_stream_key(seed, 3); \
        _Pragma("omp parallel for schedule(dynamic, 16)")                     \
        for (int64_t s = 0; s < num_systems; s++) {                           \
            int64_t o = offsets[s], e = offsets[s + 1];                       \
            for (int64_t g = o; g < e; g++) {                                 \
                double ua = u01(ka, (uint64_t)g), ub = u01(kb, (uint64_t)g);  \
                double uc = u01(kc, (uint64_t)g), ud = u01(kd, (uint64_t)g);  \
                TT av, bv, cv, dv;                                            \
                if (kind == GEN_REFGEN) {                                     \
                    av = (TT)(3.0 + ua); bv = (TT)(8.0 + ub);                 \
                    cv = (TT)(2.0 + uc); dv = (TT)(5.0 + ud);                 \
                    if (g == o) av = 0;                                       \
                    if (g == e - 1) cv = 0;                                   \
                } else if (kind == GEN_PAIRED) {                              \
                    int even = ((g - o) & 1) == 0;                            \
                    av = (TT)(even ? 0.25 * (2.0 * ua - 1.0) : 1.0 + ua);     \
                    bv = (TT)((even && g == e - 1) ? 1.0 + ub                 \
                                                   : 0.001 * (2.0 * ub - 1.0)); \
                    cv = (TT)(even ? 1.0 + uc : 0.25 * (2.0 * uc - 1.0));     \
                    dv = (TT)(2.0 * ud - 1.0);                                \
                    if (g == o) av = 0;                                       \
                    if (g == e - 1) cv = 0;                                   \
                } else if (kind == GEN_GENERAL) {                             \
                    av = (TT)(2.0 * ua - 1.0); bv = (TT)(2.0 * ub - 1.0);     \
                    cv = (TT)(2.0 * uc - 1.0); dv = (TT)(2.0 * ud - 1.0);     \
                    if (g == o) av = 0;                                       \
                    if (g == e - 1) cv = 0;                                   \
                } else {                                                      \
                    av = (TT)(2.0 * ua - 1.0);                                \
                    cv = (TT)(2.0 * uc - 1.0);                                \
                    dv = (TT)(2.0 * ud - 1.0);                                \
                    if (g == o) av = 0;                                       \
                    if (g == e - 1) cv = 0;                                   \
                    double margin = (kind == GEN_DOMINANT) ? (1.0 + ub)       \
                                                           : (0.01 + 0.09 * ub); \
                    bv = (TT)((fabs((double)av) + fabs((double)cv)) + margin); \
                }                                                             \
                a[g] = av; b[g] = bv; c[g] = cv; d[g] = dv;                   \
            }                                                                 \
        }                                                                     \
        return 0;                                                             \
    }

GEN_DRIVER(oracle_generate_f32, float)
GEN_DRIVER(oracle_generate_f64, double)
