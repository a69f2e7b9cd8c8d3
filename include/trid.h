/*
 * trid.h -- C ABI of the MI355X (gfx950) batched tridiagonal solver library
 * (libtrid.so).
 *
 * A batch is numSystems independent systems T x = d of systemSize unknowns,
 * stored system-major in five separate arrays: element j of system s is at
 * s*systemSize + j in a (sub-diagonal), b (diagonal), c (super-diagonal),
 * d (right-hand side) and x (solution). a[first] and c[last] of each system
 * are ignored (treated as 0). The reference's methods do no pivoting and are
 * meant for diagonally dominant systems; trid_gtsv_* and
 * trid_solve_pivoting_* interchange rows (LAPACK ?gtsv) for the others.
 *
 * Every function returns an int status (trid_status); none of them exits
 * the process.
 *
 * Reference: zhangyaobit/tridiagonalsolvers, directory tridiagonalsolvers/.
 * The reference declares its solver API only as forward declarations in
 * test_tridiagonal.cpp:17-26 (C++ linkage, fixed element type T, host
 * pointers, void return).
 */
#ifndef TRID_H
#define TRID_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* status, methods, options                                                 */
/* ---------------------------------------------------------------------- */

typedef enum trid_status {
    TRID_OK = 0,
    TRID_EINVAL = 1,  /* bad size, method parameter or null pointer */
    TRID_ENOMEM = 2,  /* device or host allocation failed */
    TRID_EHIP = 3,    /* a HIP runtime call or kernel launch failed */
    TRID_ENOTSUP = 4, /* valid request that the library cannot run */
    TRID_ENODEV = 5,  /* no usable gfx950 device */
    TRID_EAGAIN = 6   /* not finished / still in use: try again later */
} trid_status;

typedef enum trid_method {
    TRID_CR = 0,        /* cyclic reduction               (cr_kernel.cu)      */
    TRID_PCR = 1,       /* parallel cyclic reduction      (pcr_kernel.cu)     */
    TRID_CRPCR = 2,     /* CR down to `rest`, PCR, CR back (crpcr_kernel.cu)  */
    TRID_CRNBC = 3,     /* CR, LDS-resident, padded LDS   (crNBC.cu)          */
    TRID_CRPCRNBC = 4,  /* CR-PCR, LDS-resident, padded   (crNBC.cu)          */
    TRID_CRNBC32 = 5,   /* CR, LDS-resident, unpadded     (crNBC32banks.cu)   */
    TRID_PCRGAUSS = 6,  /* PCR to G subsystems, Thomas    (pcrGauss_*.cu)     */
    TRID_AUTO = 7       /* library chooses (plan / autotune table)            */
} trid_method;

/* Per-call options. Zero-initialise and set what you need; 0 = default. */
typedef struct trid_opts {
    int32_t rest;       /* CRPCR/CRPCRNBC: size of the system handed to PCR,
                           power of two in [2, n/2] (crpcr_app.cu:5)          */
    int32_t gaussian;   /* PCRGAUSS: number G of Thomas subsystems, power of
                           two; the PCR depth is exactly log2(G)
                           (pcrGauss_app.cu:164)                              */
    int32_t lds_pad;    /* LDS-resident kernels: one pad element per lds_pad
                           elements (power of two), -1 = no padding           */
    int32_t lanes;      /* register kernels: lanes per system (power of two,
                           2..1024; above 64 a workgroup of lanes/64 waves
                           holds one system), 0 = plan                        */
    int32_t per_lane;   /* register kernels: unknowns per lane (1..16; 2..16
                           when lanes > 64, FP64 at 16 with <= 512 lanes)     */
    int32_t io_mode;    /* register kernels, dense batches: 0 = plan,
                           1 per-lane non-temporal, 2 per-lane cached,
                           3 wave-coalesced through LDS, 4 as 3 non-temporal  */
    int32_t flags;      /* bit set of TRID_FLAG_*                             */
} trid_opts;

/* PCRGAUSS: use the two-pass form (global PCR pass, then strided Thomas pass
 * through a handle workspace of about 5*n*N elements: the normalized rows and
 * the elimination's factor pairs) even when the system fits one workgroup's
 * LDS. It is always used when it does not fit. */
#define TRID_FLAG_TWO_PASS 1

/* Human-readable text for a status code (static storage). */
const char* trid_status_string(int status);

/* Status of the most recent call made by this host thread through any entry
 * point (used by the void drop-ins of trid_compat.h). */
int trid_last_status(void);

/* Number of gfx950 devices visible, or a negative status. */
int trid_device_count(void);

/* ---------------------------------------------------------------------- */
/* Host-pointer drop-ins: one per reference wrapper                         */
/* ---------------------------------------------------------------------- */
/*
 * All pointers are host memory. Each call copies a,b,c,d to the device,
 * runs exactly one solve on device 0 (current device) and copies x back.
 * a,b,c,d are not modified. Unlike the reference wrappers there is no
 * built-in 50-launch timing loop and no printf.
 */

/* replaces void cr(T*,T*,T*,T*,T*,int,int)           cr_app.cu:4         */
int trid_cr_f32(const float* a, const float* b, const float* c, const float* d,
                float* x, int systemSize, int numSystems);
int trid_cr_f64(const double* a, const double* b, const double* c, const double* d,
                double* x, int systemSize, int numSystems);

/* replaces void pcr(T*,T*,T*,T*,T*,int,int)          pcr_app.cu:5        */
int trid_pcr_f32(const float* a, const float* b, const float* c, const float* d,
                 float* x, int systemSize, int numSystems);
int trid_pcr_f64(const double* a, const double* b, const double* c, const double* d,
                 double* x, int systemSize, int numSystems);

/* replaces void crpcr(T*,...,int,int,int restSystemSize) crpcr_app.cu:5  */
int trid_crpcr_f32(const float* a, const float* b, const float* c, const float* d,
                   float* x, int systemSize, int numSystems, int restSystemSize);
int trid_crpcr_f64(const double* a, const double* b, const double* c, const double* d,
                   double* x, int systemSize, int numSystems, int restSystemSize);

/* replaces void crNBC(T*,T*,T*,T*,T*,int,int)        crNBC.cu:123        */
int trid_crnbc_f32(const float* a, const float* b, const float* c, const float* d,
                   float* x, int systemSize, int numSystems);
int trid_crnbc_f64(const double* a, const double* b, const double* c, const double* d,
                   double* x, int systemSize, int numSystems);

/* replaces void crpcrNBC(T*,...,int,int,int restSystemSize) crNBC.cu:344 */
int trid_crpcrnbc_f32(const float* a, const float* b, const float* c, const float* d,
                      float* x, int systemSize, int numSystems, int restSystemSize);
int trid_crpcrnbc_f64(const double* a, const double* b, const double* c, const double* d,
                      double* x, int systemSize, int numSystems, int restSystemSize);

/* replaces void crNBC32(T*,T*,T*,T*,T*,int,int)      crNBC32banks.cu:123 */
int trid_crnbc32_f32(const float* a, const float* b, const float* c, const float* d,
                     float* x, int systemSize, int numSystems);
int trid_crnbc32_f64(const double* a, const double* b, const double* c, const double* d,
                     double* x, int systemSize, int numSystems);

/* replaces void pcrGaussWrapper(T*,...,int,int,int gaussian) pcrGauss_app.cu:164 */
int trid_pcrgauss_f32(const float* a, const float* b, const float* c, const float* d,
                      float* x, int systemSize, int numSystems, int gaussian);
int trid_pcrgauss_f64(const double* a, const double* b, const double* c, const double* d,
                      double* x, int systemSize, int numSystems, int gaussian);

/* replaces void tridiagonalMKL(T*,T*,T*,T*&,T*&,int,int) tridiagonalMKL.cpp:5
 * (MKL ?gtsv per system, partial pivoting). Host pointers; a, b, c, d are not
 * modified and x receives the solution (the reference overwrites d and swaps
 * the d/x pointers). A singular system's x is NaN. systemSize >= 1.       */
int trid_gtsv_f32(const float* a, const float* b, const float* c, const float* d,
                  float* x, int systemSize, int numSystems);
int trid_gtsv_f64(const double* a, const double* b, const double* c, const double* d,
                  double* x, int systemSize, int numSystems);

/* ---------------------------------------------------------------------- */
/* Device-pointer, stream-ordered API                                       */
/* ---------------------------------------------------------------------- */
/*
 * No counterpart in the reference (its wrappers allocate and copy on every
 * call, cr_app.cu:20-65). All array pointers are device memory; `stream` is
 * a hipStream_t (NULL = default stream). The solve is enqueued and the call
 * returns without synchronising. A handle owns the workspace some methods
 * need; use one handle from one host thread at a time.
 */
typedef struct trid_context* trid_handle;

int trid_create(trid_handle* handle);
int trid_destroy(trid_handle handle);

int trid_solve_f32(trid_handle handle, int method,
                   const float* a, const float* b, const float* c, const float* d,
                   float* x, int64_t systemSize, int64_t numSystems,
                   const trid_opts* opts, void* stream);
int trid_solve_f64(trid_handle handle, int method,
                   const double* a, const double* b, const double* c, const double* d,
                   double* x, int64_t systemSize, int64_t numSystems,
                   const trid_opts* opts, void* stream);

/*
 * Solve without device allocations. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). It is
 * trid_solve_f32 (elemSize 4, float arrays) or trid_solve_f64 (elemSize 8,
 * double arrays) without a handle: same methods, options, kernels and
 * results. The workspace some plans need (cr / crpcr by separators past the
 * register sizes, the two-pass pcrGauss form) is one fixed device array of
 * TRID_NOALLOC_SCRATCH_BYTES per device, part of the library and held until
 * the process ends; a request that needs more returns TRID_ENOMEM before
 * anything is enqueued (64 MiB holds the separator records of 64 systems of
 * 2^20 unknowns in FP32 or FP64, and the two-pass pcrGauss workspace of
 * 3.3e6 FP32 / 1.6e6 FP64 unknowns). The call allocates nothing; its first call on
 * a device looks up that device's array address. Every call on one device
 * shares the array and finds the previous call's contents (no kernel here
 * needs it cleared): calls that overlap on one device (different streams or
 * threads) race on it and give wrong results, so order them on one stream.
 */
#define TRID_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int trid_solve_noalloc(int method, int elemSize, const void* a, const void* b, const void* c,
                       const void* d, void* x, int64_t systemSize, int64_t numSystems,
                       const trid_opts* opts, void* stream);

/*
 * FP32 solve with FP64 iterative refinement. No counterpart in the reference
 * (its element type is fixed at build time, common.h:5-12). a,b,c,d,x are
 * FP64 device arrays; the solves run in FP32 with `method` (TRID_CR,
 * TRID_CRPCR, TRID_PCR or TRID_AUTO); then `iters` times (0..8):
 * r = d - T x in FP64, solve T e = r in FP32, x += e in FP64. One kernel per
 * call; x is written once. Systems up to 8192 unknowns.
 */
int trid_solve_refined_f64(trid_handle handle, int method,
                           const double* a, const double* b, const double* c, const double* d,
                           double* x, int64_t systemSize, int64_t numSystems, int iters,
                           const trid_opts* opts, void* stream);

/*
 * Strided batch: system s occupies [s * batchStride, s * batchStride +
 * systemSize) of each array (batchStride >= systemSize), the layout of
 * rocSPARSE's gtsv_no_pivot_strided_batch; batchStride == systemSize is the
 * contiguous batch of trid_solve_*. No counterpart in the reference
 * (contiguous only). Methods TRID_CR, TRID_PCR, TRID_CRPCR and TRID_AUTO on
 * the register kernels (systemSize up to 16384 FP32 / 8192 FP64); CR and
 * CR-PCR above that by separators (handle workspace, as trid_solve_*); other
 * methods or sizes return TRID_ENOTSUP. Elements between systems are neither
 * read nor written.
 */
int trid_solve_strided_f32(trid_handle handle, int method, const float* a, const float* b,
                           const float* c, const float* d, float* x, int64_t systemSize,
                           int64_t numSystems, int64_t batchStride, const trid_opts* opts,
                           void* stream);
int trid_solve_strided_f64(trid_handle handle, int method, const double* a, const double* b,
                           const double* c, const double* d, double* x, int64_t systemSize,
                           int64_t numSystems, int64_t batchStride, const trid_opts* opts,
                           void* stream);

/*
 * Interleaved batch: row i of system j at i * batchStride + j (batchStride >=
 * numSystems), the layout of rocSPARSE's gtsv_interleaved_batch. No
 * counterpart in the reference (system-major only). Systems of 257..1024
 * (FP32) or 129..1024 (FP64) unknowns are gathered through LDS, a few per
 * workgroup, and solved by the register CR-PCR kernel (one wave each), with
 * no workspace; other sizes by Gaussian elimination from both ends at once
 * (two lanes per system, every access coalesced across systems), which keeps
 * its factors in the handle's workspace (systemSize * batchStride elements),
 * so such calls on one handle must not overlap on different streams. The a
 * of row 0 and the c of the last row are ignored; x may be d (in place) but
 * no other input.
 */
int trid_solve_interleaved_f32(trid_handle handle, const float* a, const float* b,
                               const float* c, const float* d, float* x, int64_t systemSize,
                               int64_t numSystems, int64_t batchStride, void* stream);
int trid_solve_interleaved_f64(trid_handle handle, const double* a, const double* b,
                               const double* c, const double* d, double* x, int64_t systemSize,
                               int64_t numSystems, int64_t batchStride, void* stream);

/*
 * Partial-pivoting solve: Gaussian elimination with row interchanges, the
 * LAPACK ?gtsv algorithm for one right-hand side, for systems that are not
 * diagonally dominant. The device-side counterpart of the reference's MKL
 * baseline (tridiagonalMKL.cpp:5-26, ?gtsv per system). Any systemSize >= 1.
 * `info` (device array of numSystems int32, may be NULL) receives 0, or the
 * 1-based row of the first exactly zero pivot (LAPACK's INFO); such a system
 * is singular and its x is set to NaN. a, b, c, d are not modified; x may be
 * d but no other input. Uses the handle's workspace (a checkpoint of 4
 * elements per 128 bytes of each system), so calls on one handle must not
 * overlap on different streams. The interleaved form takes row i of system s
 * at i * batchStride + s (batchStride >= numSystems), the layout of
 * rocSPARSE's gtsv_interleaved_batch.
 */
int trid_solve_pivoting_f32(trid_handle handle, const float* a, const float* b, const float* c,
                            const float* d, float* x, int64_t systemSize, int64_t numSystems,
                            int32_t* info, void* stream);
int trid_solve_pivoting_f64(trid_handle handle, const double* a, const double* b,
                            const double* c, const double* d, double* x, int64_t systemSize,
                            int64_t numSystems, int32_t* info, void* stream);
int trid_solve_pivoting_interleaved_f32(trid_handle handle, const float* a, const float* b,
                                        const float* c, const float* d, float* x,
                                        int64_t systemSize, int64_t numSystems,
                                        int64_t batchStride, int32_t* info, void* stream);
int trid_solve_pivoting_interleaved_f64(trid_handle handle, const double* a, const double* b,
                                        const double* c, const double* d, double* x,
                                        int64_t systemSize, int64_t numSystems,
                                        int64_t batchStride, int32_t* info, void* stream);

/*
 * Relative residual of every system, in FP64 on the device:
 * relres[s] = max_i |(d - T x)_i| / max_i |d_i| over system s =
 * [offsets[s], offsets[s+1]) (device arrays; a[first], c[last] ignored).
 * No counterpart in the reference (its check is an RMS against the CPU
 * solution, tridiagonal_gold.h:97-148).
 */
int trid_residual_f64(const double* a, const double* b, const double* c, const double* d,
                      const double* x, const int64_t* offsets, int64_t numSystems,
                      double* relres, void* stream);

/*
 * Seeded synthetic batch written straight into device memory (same values as
 * the reference-shaped generator kinds below, bit for bit, on every run):
 *   kind 0  a,c ~ U[-1,1], b = |a|+|c|+U[1,2],      d ~ U[-1,1]
 *   kind 1  a,c ~ U[-1,1], b = |a|+|c|+U[0.01,0.1], d ~ U[-1,1]
 *   kind 2  b = 8+U, a = 3+U, c = 2+U, d = 5+U (tridiagonal_gold.h:81-94)
 *   kind 3  not diagonally dominant but well conditioned once rows are
 *           interchanged: c = 1+U on even rows and a = 1+U on odd rows
 *           couple each row pair, b = 0.001 U[-1,1] (b = 1+U on the last row
 *           of an odd-sized system), the other couplings 0.25 U[-1,1],
 *           d ~ U[-1,1]; the no-pivoting methods meet tiny pivots on it
 *   kind 4  a, b, c, d ~ U[-1,1] (general; may be ill-conditioned)
 * with a[first] = c[last] = 0 in every system (U = U[0,1]).
 */
int trid_generate_f32(int kind, uint64_t seed, float* a, float* b, float* c, float* d,
                      int64_t systemSize, int64_t numSystems, void* stream);
int trid_generate_f64(int kind, uint64_t seed, double* a, double* b, double* c, double* d,
                      int64_t systemSize, int64_t numSystems, void* stream);

/* ---------------------------------------------------------------------- */
/* Ragged batches: one persistent, graph-captured launch                    */
/* ---------------------------------------------------------------------- */
/*
 * No counterpart in the reference (one systemSize per call). System s of a
 * ragged batch occupies [offsets[s], offsets[s+1]) of flat FP64 device
 * arrays a, b, c, d, x; sizes 2..16384. A plan buckets the systems by padded
 * size (power of two) and builds the work queue of one persistent kernel
 * (workgroup items for 1024..16384, wave items below); each bucket solves in
 * FP32 with its (per_lane, lanes, rest) configuration and refines in FP64
 * `iters` times, or (TRID_RAGGED_FP64) solves once in FP64 with the same
 * configuration. The plan depends only on the sizes and the table.
 */
typedef struct trid_bucket_cfg {
    int32_t size;      /* padded system size (power of two, 2..16384)           */
    int32_t per_lane;  /* unknowns per lane: <= 8 (16 for sizes 8192, 16384)   */
    int32_t lanes;     /* lanes per system: size/per_lane (<= 64 below 1024)    */
    int32_t rest;      /* CR -> PCR switch size, power of two <= lanes          */
} trid_bucket_cfg;

typedef struct trid_ragged_info {
    int64_t num_systems;
    int64_t total_unknowns;
    int32_t num_buckets;
    int32_t workgroup_items;
    int32_t wave_items;
    int32_t grid;        /* persistent workgroups launched */
} trid_ragged_info;

typedef struct trid_ragged_plan_s* trid_ragged_plan;

/* Do not capture/replay a hipGraph (a plain launch); also used when stream is
 * NULL, which cannot be captured. */
#define TRID_RAGGED_NO_GRAPH 1
/* Solve every system in one FP64 pass of the same kernels instead of FP32 +
 * FP64 refinement (`iters` must be 0). One read of a, b, c, d and one write
 * of x per unknown; systems up to 8192 unknowns (TRID_ENOTSUP if the plan
 * holds larger ones). */
#define TRID_RAGGED_FP64 2

/* Reads an autotune table file: the JSON that tridiagonalsolvers_amd.autotune
 * writes (tune/ragged_f64_refine.json, tune/ragged_f64_fp64.json), an object
 * whose "buckets" array holds objects with integer "size", "per_lane",
 * "lanes" and "rest" (other keys are skipped). The first `capacity` buckets
 * go to table[], *count receives how many the file holds. Each entry is
 * checked as trid_ragged_plan_create checks it. TRID_EINVAL for a file that
 * cannot be read, malformed JSON, a bucket without one of the four keys or
 * an invalid configuration. Host only (no device needed). */
int trid_bucket_table_load(const char* path, trid_bucket_cfg* table, int capacity, int* count);

/* offsets: host array of numSystems+1 element offsets. table: optional
 * per-bucket configurations (e.g. from the autotuner, or read by
 * trid_bucket_table_load); buckets not listed use the library default. */
int trid_ragged_plan_create(trid_ragged_plan* plan, const int64_t* offsets, int64_t numSystems,
                            const trid_bucket_cfg* table, int tableSize);
int trid_ragged_plan_destroy(trid_ragged_plan plan);
int trid_ragged_plan_info(trid_ragged_plan plan, trid_ragged_info* info,
                          trid_bucket_cfg* buckets, int64_t* bucketSystems, int capacity);
/* Enqueue the ragged solve on `stream`. The first call for a given set of
 * pointers, `iters` and mode captures the reset + persistent launch into a
 * hipGraph; later calls replay it. */
int trid_ragged_solve_f64(trid_ragged_plan plan, const double* a, const double* b,
                          const double* c, const double* d, double* x, int iters, int flags,
                          void* stream);
/* The refine mode to a residual target: after the FP32 solve a system takes
 * FP64 refinements until max_i |d_i - (A x)_i| <= tol * max_i |d_i| (its
 * relative residual, the FP64 residual of the current x; max|r| is taken
 * from the residual rounded to FP32 and scaled by 1 + 2^-23, an upper bound)
 * or it has taken maxIters (0..8) of them. The decision is made per system,
 * so a system's result does not depend on the other systems of the batch;
 * a workgroup stops once all of its systems have. tol = 0: exactly maxIters
 * refinements (trid_ragged_solve_f64). TRID_EINVAL for a negative or
 * non-finite tol, or with TRID_RAGGED_FP64. */
int trid_ragged_solve_tol_f64(trid_ragged_plan plan, const double* a, const double* b,
                              const double* c, const double* d, double* x, int maxIters,
                              double tol, int flags, void* stream);

/* ---------------------------------------------------------------------- */
/* Resident ADI grids (2D heat equation, Peaceman-Rachford)                 */
/* ---------------------------------------------------------------------- */
/*
 * No counterpart in the reference. An n x n FP64 grid u (row-major, zero
 * Dirichlet boundary) stays on the device; each step with parameter r solves
 * (I - r dxx) u* = (I + r dyy) u   (n row systems tridiag(-r, 1+2r, -r))
 * (I - r dyy) u' = (I + r dxx) u*  (n column systems, stride n)
 * with the LU factors of tridiag(-r, 1+2r, -r), computed once per call and
 * resident on the device, applied to the rows and, on the transposed grid
 * (tiled transposes through LDS), to the columns.
 */
typedef struct trid_adi_s* trid_adi;

/* u_device: an existing n*n device grid to step in place, or NULL to
 * allocate one (trid_adi_grid returns it). 2 <= n <= 16384 (TRID_EINVAL
 * otherwise). */
int trid_adi_create(trid_adi* adi, int64_t n, double* u_device);
int trid_adi_destroy(trid_adi adi);
double* trid_adi_grid(trid_adi adi);
/* Enqueue `steps` steps with parameter r > 0 on `stream`. */
int trid_adi_step(trid_adi adi, double r, int steps, void* stream);

/* ---------------------------------------------------------------------- */
/* Continuous-batching request server                                       */
/* ---------------------------------------------------------------------- */
/*
 * No counterpart in the reference (a single-threaded test driver). Requests
 * (one or more FP64 systems, sizes 2..16384, host arrays) are submitted from
 * any thread; a scheduler thread admits them by device-memory footprint,
 * batches what has arrived into ragged solves (FP32 + FP64 refinement) on
 * `slots` streams, overlaps batches on the device, and copies results back.
 * Registered operators keep coefficient arrays resident; requests on them
 * send only d. Each request's result is independent of the batch it lands in.
 */
typedef struct trid_server_s* trid_server;

/* Fake executor: no device work at all, nothing is computed; batches finish
 * after a modelled delay. For testing admission and scheduling only. */
#define TRID_SERVER_FAKE_EXECUTOR 1

typedef struct trid_server_opts {
    int64_t hbm_budget_bytes;    /* device memory reserved for batches and operators (0 = 4 GiB) */
    int32_t slots;               /* batches in flight / streams (0 = 3)                  */
    int32_t max_batch_systems;   /* (0 = 65536)                                           */
    int64_t max_batch_unknowns;  /* (0 = 2^23)                                            */
    int32_t batch_window_us;     /* while the device is busy, wait up to this long for a batch to grow (0 = 200) */
    int32_t iters;               /* FP64 refinement steps (0..8; -1 = 2)                  */
    int32_t flags;               /* TRID_SERVER_*                                         */
    int32_t fake_fail_every;     /* fake executor: fail every k-th batch (0 = never)      */
    int32_t fake_batch_us;       /* fake executor: time a batch takes (0 = size model)    */
    const trid_bucket_cfg* bucket_table;  /* optional per-bucket configurations          */
    int32_t bucket_table_size;
} trid_server_opts;

#define TRID_REQ_SOLVE 0
#define TRID_REQ_OPERATOR 1
#define TRID_REQ_ADI 2

typedef struct trid_request {
    int32_t kind;                /* TRID_REQ_SOLVE, TRID_REQ_OPERATOR or TRID_REQ_ADI     */
    int64_t num_systems;         /* SOLVE: systems in the request                         */
    const int64_t* offsets;      /* SOLVE: host, num_systems+1, into a,b,c,d,x            */
    const double* a;             /* SOLVE: host coefficients                              */
    const double* b;
    const double* c;
    const double* d;             /* SOLVE, OPERATOR: host right-hand sides                */
    double* x;                   /* SOLVE, OPERATOR: host solution, written before the
                                    request completes                                     */
    int32_t op;                  /* OPERATOR: resident operator id                        */
    int32_t adi;                 /* ADI: resident grid id, advanced by adi_steps
                                    Peaceman-Rachford steps with parameter adi_r          */
    int32_t adi_steps;
    double adi_r;
} trid_request;

typedef struct trid_request_result {
    int32_t status;              /* TRID_OK, or why the request failed                    */
    int32_t state;               /* 2 done, 3 cancelled, 4 failed                         */
    double latency_us;           /* submit -> results copied back                         */
    double queue_us;             /* submit -> batch launched                              */
    int64_t batch;               /* batch id it ran in                                    */
    int64_t batch_systems;       /* systems in that batch                                 */
} trid_request_result;

typedef struct trid_server_stats {
    int64_t submitted, completed, failed, cancelled, batches;
    int64_t queued;
    int64_t device_bytes_in_use;  /* batches in flight + resident operators              */
    int64_t peak_device_bytes;
    int64_t budget_bytes;
} trid_server_stats;

int trid_server_create(trid_server* server, const trid_server_opts* opts);
/* Finishes every queued and running request, then frees everything. */
int trid_server_destroy(trid_server server);
/* sizes: numSystems system sizes; a, b, c: host arrays of their sum. */
int trid_server_register_operator(trid_server server, const int64_t* sizes, int64_t numSystems,
                                  const double* a, const double* b, const double* c,
                                  int32_t* op);
/* TRID_EAGAIN while requests on the operator are queued or running. */
int trid_server_release_operator(trid_server server, int32_t op);
/* A resident n x n FP64 ADI grid (initial values u0, host) inside the
 * server's budget; requests with .adi = id step it. Steps on one grid run in
 * submission order, never two at once. */
int trid_server_create_adi(trid_server server, int64_t n, const double* u0, int32_t* adi);
/* Copy the grid to the host; TRID_EAGAIN while requests on it are queued or
 * running. */
int trid_server_read_adi(trid_server server, int32_t adi, double* u);
int trid_server_release_adi(trid_server server, int32_t adi);
/* TRID_ENOMEM if the request can never fit the budget or the batch limits.
 * A batch occupies one contiguous block of the budget, and resident
 * operators and ADI grids are placed from the top of it; a request that fits
 * no free block while no solve batch is in flight (so that no retirement can
 * free memory) is not kept waiting: it completes as failed with status
 * TRID_ENOMEM. The caller keeps every array alive until trid_server_wait
 * reports it. */
int trid_server_submit(trid_server server, const trid_request* request, uint64_t* ticket);
/* timeout_us < 0 waits without limit; TRID_EAGAIN on timeout. After a final
 * state is reported the ticket is retired (later waits return TRID_EINVAL);
 * threads already waiting on the same ticket all receive the result. */
int trid_server_wait(trid_server server, uint64_t ticket, int64_t timeout_us,
                     trid_request_result* result);
/* Page-locked host memory for request arrays (a, b, c, d, x): arrays that lie
 * inside such a buffer are copied to and from the device by DMA directly,
 * with no staging copy. Free a buffer only after every request that uses it
 * has been waited for; trid_server_destroy frees the remaining ones. */
int trid_server_host_alloc(trid_server server, size_t bytes, void** ptr);
int trid_server_host_free(trid_server server, void* ptr);
/* Cancels a request that is still queued; TRID_EINVAL once it is running. */
int trid_server_cancel(trid_server server, uint64_t ticket);
int trid_server_get_stats(trid_server server, trid_server_stats* stats);

/* ---------------------------------------------------------------------- */
/* Plan introspection (host only; no device needed)                         */
/* ---------------------------------------------------------------------- */
/*
 * No counterpart in the reference, whose launch shapes are implied by
 * systemSize (one block of n/2 or n threads per system, cr_app.cu:37-38).
 * Validates a request exactly as trid_solve_* would and reports the kernel
 * family and launch shape it would use: family 0 = register-resident
 * CR/CR-PCR/PCR, 1 = LDS-resident CR/CR-PCR with pad map, 2 = PCR + Thomas,
 * 3 = CR/CR-PCR by separators (systems longer than 16,384 unknowns in FP32
 * or 8,192 in FP64: chunks of 1,024 rows with per_lane 16, lanes 64).
 */
typedef struct trid_plan_info {
    int32_t family;
    int32_t per_lane;    /* family 0: unknowns per lane             */
    int32_t lanes;       /* family 0: lanes per system              */
    int32_t rest;        /* families 0/1: PCR system size (2 = CR)  */
    int32_t lds_pad;     /* family 1: pad period, 0 = unpadded      */
    int32_t gaussian;    /* family 2: Thomas subsystems             */
    int64_t padded_size; /* systemSize rounded up to a power of two */
} trid_plan_info;

int trid_plan_query(int method, int64_t systemSize, int64_t numSystems, int elemSize,
                    const trid_opts* opts, trid_plan_info* out);

#ifdef __cplusplus
}
#endif

#endif /* TRID_H */
