// kern_ragged.hip -- persistent launches over a ragged batch.
//
// No counterpart in the reference (every reference wrapper takes one
// systemSize for the whole batch). A ragged batch is numSystems systems of
// arbitrary sizes n_s >= 2 in flat FP64 arrays, system s occupying
// [offsets[s], offsets[s+1]). Every system is solved in FP32 on the register
// CR-PCR core and refined in FP64 (the kern_refine.hip scheme), or solved once
// in FP64, padded to the power of two above n_s with identity rows (exact).
//
// Work queue (built on the host, trid_ragged.cpp): systems are bucketed by
// padded size; each bucket has a configuration (unknowns per lane M, lanes
// per system, CR->PCR switch size) chosen by the autotune table. Four item
// classes, each run by a persistent kernel of its own (the four go into one
// captured graph, so that each gets the register and LDS budget its shape
// needs):
//   * workgroup items, 1024..8192 unknowns, drained from atomic queue
//     counters: W waves per system, 8/W systems to an item of a 512-thread
//     workgroup (one system per group of W waves: barriers are
//     workgroup-wide, and every group of an item runs the same bucket's
//     code, so they match). At 16 unknowns per lane (8192 unknowns, and the
//     buckets below that the table sets so) ragged_wg_kernel runs them with
//     256 registers, one workgroup per CU; at 8 per lane ragged_half_kernel
//     runs them with 128, two workgroups per CU: the solve's reduced-system
//     phase is a chain of LDS round trips and barriers, and the second
//     workgroup works while the first waits;
//   * wide items, 8193..16384 unknowns, refine mode only
//     (ragged_wide_kernel): one system per 1024-thread workgroup at 16
//     unknowns per lane, the coefficients read again on each pass;
//   * wave items, up to 512 unknowns (ragged_wave_kernel): up to 64/L
//     systems of one bucket per item, solved by one wave, L lanes each (no
//     barrier); 256-thread workgroups, items dealt round-robin over the waves.
// The refine mode reads each system's FP64 coefficients once: b and d stay in
// registers and a and c in LDS across its passes. A small kernel resets the
// queue counters before every launch (so the whole sequence can be captured
// in a hipGraph and replayed).
#include "reg_core.hpp"
#include "trid_device.hpp"
#include "trid_launch.hpp"

#include <mutex>
#include <type_traits>

namespace trid {

namespace {

// Coefficient arrays whose loads are in flight at once in ragged_rows (the
// wide items' passes): 2, or sequential above 8 unknowns per lane.
#ifndef TRID_RAGGED_AHEAD
#define TRID_RAGGED_AHEAD 2
#endif

constexpr int kWgThreads = 1024;    // wide items: 16 waves, 16 unknowns per lane
constexpr int kWgWaves = kWgThreads / kWave;
static_assert(kWgWaves == 16, "wide items assume 16 waves");
constexpr int kFullThreads = 512;   // 8192-unknown items: 8 waves, 16 unknowns per lane
constexpr int kFullWaves = kFullThreads / kWave;
constexpr int kHalfThreads = 512;   // workgroup items of at most 8 waves, 8 per lane
constexpr int kHalfWaves = kHalfThreads / kWave;
constexpr int kWaveThreads = 256;   // wave-item kernel: 4 independent waves
constexpr int kWaveWaves = kWaveThreads / kWave;

// Transposition area of one wave (coalesced loads and stores, below).
// Row r holds the M consecutive unknowns of lane r, padded by 16 bytes:
// a row stride of M*8+16 bytes puts the 16 lanes of one pass of a 16-byte
// LDS access on disjoint banks.
template <int M>
struct RStage {
    static constexpr int kRowBytes = M * 8 + 16;
    static constexpr int kRounds = M > 8 ? M / 8 : 1;
    static constexpr int kQ = M / kRounds;         // loads per lane per round
    static constexpr int kRows = kWave / kRounds;  // rows per round
    static constexpr int kBytes = kRows * kRowBytes;
};
constexpr int kStageWaveBytes = RStage<8>::kBytes;  // 5120

// Kernel modes: FP32 solve + FP64 refinement, or one FP64 solve (systems up
// to 8192 unknowns).
// kRefineTol: the refine mode with the residual check (a system stops
// refining at the target); a kernel of its own, so that the fixed-count
// kernels keep their registers.
enum RaggedMode { kRefine = 0, kDirect = 1, kRefineTol = 2 };

// The refine mode keeps each system's FP64 a and c in LDS across its
// passes (and b, d in registers), so the coefficients are read from HBM once:
// every wave owns a home area of 64 rows of RM doubles (RM = 8, or 16 for
// 16 unknowns per lane) for each of a and c, row r holding the operator rows
// of lane r. A row's 16-byte chunks are XOR-permuted by bits of the row index
// (bits 2..3 for the four chunks of an 8-double row, 1..3 for the eight of a
// 16-double row), so that 16 lanes reading one chunk of their own rows (rows
// 0..15) touch 16 distinct groups of four of the 64 banks, and 32 lanes
// writing whole consecutive rows do too. The home area also transposes b, d
// (on their way to the registers) and x.
template <int RM>
struct HomeT {
    static_assert(RM == 8 || RM == 16, "home rows of 8 or 16 doubles");
    static constexpr int kRowBytes = RM * 8;
    static constexpr int kArrBytes = kWave * kRowBytes;  // 4 / 8 KiB
    static constexpr int kWaveBytes = 2 * kArrBytes;     // a, then c
    __device__ static __forceinline__ int at(int r, int k)
    {
        const int perm = RM == 8 ? ((r >> 2) & 3) : ((r >> 1) & 7);
        return r * kRowBytes + (((k >> 1) ^ perm) << 4) + ((k & 1) << 3);
    }
};
// home layout of M unknowns per lane
template <int M>
using Home = HomeT<(M > 8 ? 16 : 8)>;

// Per-wave area (home area or transposition area) of a mode at M unknowns
// per lane.
template <int MODE, int M>
constexpr size_t ragged_wave_bytes()
{
    return MODE != kDirect ? (size_t)Home<M>::kWaveBytes
                           : (size_t)(M > 8 ? RStage<M>::kBytes : kStageWaveBytes);
}
// LDS of one group of W waves (one system): the exchange area of wg_core,
// then 2 W boundary unknowns (the first and last unknown of every wave).
template <typename T, int M>
constexpr size_t ragged_group_bytes(int W)
{
    return wg_core_lds_bytes<T, M>(W) + (size_t)2 * W * sizeof(double);
}
// Exchange areas of the workgroup items of a WAVES-wave workgroup at M
// unknowns per lane: WAVES/W groups of W waves (ragged_group_bytes is linear
// in W, so one group of all the waves is the largest).
template <int MODE, int WAVES, int M>
constexpr size_t ragged_xchg_bytes()
{
    return MODE == kDirect ? ragged_group_bytes<double, M>(WAVES)
                           : ragged_group_bytes<float, M>(WAVES);
}
template <int MODE, int WAVES, int M>
constexpr size_t wg_lds_bytes()
{
    return (size_t)WAVES * ragged_wave_bytes<MODE, M>() + ragged_xchg_bytes<MODE, WAVES, M>();
}
// Other LDS of the workgroup-item kernels: the item slots and the bucket table.
constexpr size_t kWgSmallLds = 2 * sizeof(int) + kRaggedMaxBuckets * sizeof(RaggedBucket);
// wide items: the transposition areas (16 unknowns per lane), then the
// exchange area and 2 W boundary unknowns
constexpr size_t kWideStageBytes = (size_t)kWgWaves * RStage<16>::kBytes;
constexpr size_t kWideLdsBytes =
    kWideStageBytes + ragged_group_bytes<float, 16>(16);
static_assert(wg_lds_bytes<kDirect, kFullWaves, 16>() + kWgSmallLds <= 160 * 1024,
              "LDS of the FP64 8192-unknown items");
static_assert(wg_lds_bytes<kRefine, kFullWaves, 16>() + kWgSmallLds <= 160 * 1024,
              "LDS of the refine 8192-unknown items");
// two 512-thread workgroups per CU (LDS is allocated in 512-byte blocks)
constexpr size_t lds_blocks(size_t v) { return (v + 511) / 512 * 512; }
static_assert(2 * lds_blocks(wg_lds_bytes<kDirect, kHalfWaves, 8>() + kWgSmallLds) <=
                  160 * 1024,
              "two FP64 8-wave workgroups per CU");
static_assert(2 * lds_blocks(wg_lds_bytes<kRefine, kHalfWaves, 8>() + kWgSmallLds) <=
                  160 * 1024,
              "two refine 8-wave workgroups per CU");
static_assert(kWideLdsBytes <= 160 * 1024 - 512, "LDS of the wide items");

// Lane-blocked gathers. The system (segment of L lanes, first row of the
// segment at lane - ll) starts at `base` and has n unknowns; lane ll must
// end with unknowns [ll*M, ll*M+M) (`fill` past n). Load q of lane ll reads
// unknown q*L+ll (fetch_cols), so every load instruction of the wave reads L
// consecutive doubles of each system (coalesced whatever the alignment of
// `base`); the LDS area then transposes them into rows (transpose_rows).
// Fetching all four arrays before the first transpose keeps one memory
// round trip per pass instead of four.
template <int M>
__device__ __forceinline__ void fetch_cols(const double* __restrict__ X, int64_t base, int n,
                                           int L, int ll, bool live, double fill, double (&t)[M])
{
#pragma unroll
    for (int q = 0; q < M; ++q) {
        const int j = q * L + ll;
        t[q] = (live && j < n) ? X[base + j] : fill;
    }
}

template <int M>
__device__ __forceinline__ void transpose_rows(const double (&t)[M], int L, int ll, int lane,
                                               char* stage, double (&v)[M])
{
    if constexpr (M == 1) {
        v[0] = t[0];
    } else {
        using S = RStage<M>;
        const int seg_row = lane - ll;
#pragma unroll
        for (int h = 0; h < S::kRounds; ++h) {
            lds_poison_wave(stage, S::kBytes, lane);
#pragma unroll
            for (int q = 0; q < S::kQ; ++q) {
                const int j = (h * S::kQ + q) * L + ll;
                const int row = seg_row + j / M - h * S::kRows;
                *reinterpret_cast<double*>(stage + row * S::kRowBytes + (j % M) * 8) =
                    t[h * S::kQ + q];
            }
            __builtin_amdgcn_wave_barrier();
            if (lane / S::kRows == h) {
                const char* src = stage + (lane - h * S::kRows) * S::kRowBytes;
#pragma unroll
                for (int k = 0; k < M / 2; ++k) {
                    const double2 w = *reinterpret_cast<const double2*>(src + 16 * k);
                    v[2 * k] = w.x;
                    v[2 * k + 1] = w.y;
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// The four coefficient arrays of the lane's rows, lane-blocked. Up to 8
// unknowns per lane all loads are issued before the first transpose; at 16
// (the 16384-unknown systems) the arrays go one at a time to bound the
// registers.
template <int M>
__device__ __forceinline__ void gather4(const RaggedArgs& g, int64_t base, int n, int L,
                                        int ll, int lane, bool live, char* stage,
                                        double (&va)[M], double (&vb)[M], double (&vc)[M],
                                        double (&vd)[M])
{
    // The index arithmetic is cheap; without this the compiler hoists every
    // address of every gather out of the refinement loop and spills them.
    asm volatile("" : "+v"(ll), "+v"(n));
    if constexpr (M <= 8) {
        double ta[M], tb[M], tc[M], td[M];
        fetch_cols<M>(g.d, base, n, L, ll, live, 0.0, td);
        fetch_cols<M>(g.b, base, n, L, ll, live, 1.0, tb);
        fetch_cols<M>(g.a, base, n, L, ll, live, 0.0, ta);
        fetch_cols<M>(g.c, base, n, L, ll, live, 0.0, tc);
        transpose_rows<M>(td, L, ll, lane, stage, vd);
        transpose_rows<M>(tb, L, ll, lane, stage, vb);
        transpose_rows<M>(ta, L, ll, lane, stage, va);
        transpose_rows<M>(tc, L, ll, lane, stage, vc);
    } else {
        double t[M];
        fetch_cols<M>(g.d, base, n, L, ll, live, 0.0, t);
        transpose_rows<M>(t, L, ll, lane, stage, vd);
        fetch_cols<M>(g.b, base, n, L, ll, live, 1.0, t);
        transpose_rows<M>(t, L, ll, lane, stage, vb);
        fetch_cols<M>(g.a, base, n, L, ll, live, 0.0, t);
        transpose_rows<M>(t, L, ll, lane, stage, va);
        fetch_cols<M>(g.c, base, n, L, ll, live, 0.0, t);
        transpose_rows<M>(t, L, ll, lane, stage, vc);
    }
}

// Inverse of gather_rows: the lane's M unknowns go back to unknowns
// [ll*M, ll*M+M) of the system through the LDS area, written with
// coalesced stores (nothing past n).
template <int M>
__device__ __forceinline__ void scatter_rows(double* __restrict__ X, int64_t base, int n,
                                             int L, int ll, int lane, bool live, char* stage,
                                             const double (&v)[M])
{
    if constexpr (M == 1) {
        if (live && ll < n) X[base + ll] = v[0];
    } else {
        using S = RStage<M>;
        const int seg_row = lane - ll;
        asm volatile("" : "+v"(ll), "+v"(n));
#pragma unroll
        for (int h = 0; h < S::kRounds; ++h) {
            lds_poison_wave(stage, S::kBytes, lane);
            if (lane / S::kRows == h) {
                char* dst = stage + (lane - h * S::kRows) * S::kRowBytes;
#pragma unroll
                for (int k = 0; k < M / 2; ++k)
                    *reinterpret_cast<double2*>(dst + 16 * k) = double2{v[2 * k], v[2 * k + 1]};
            }
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int q = 0; q < S::kQ; ++q) {
                const int j = (h * S::kQ + q) * L + ll;
                const int row = seg_row + j / M - h * S::kRows;
                const double t =
                    *reinterpret_cast<const double*>(stage + row * S::kRowBytes + (j % M) * 8);
                if (live && j < n) X[base + j] = t;
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// FP64 operator of the lane's unknowns [j0, j0+M) (arguments as
// ragged_rows). Identity rows past n; a[first] and c[last] ignored.
template <int M>
__device__ __forceinline__ void direct_rows(const RaggedArgs& g, int64_t base, int nb,
                                            int j0, int n, int L, int ll, int lane,
                                            bool live, char* stage, Eq<double> (&e)[M])
{
    double va[M], vb[M], vc[M], vd[M];
    gather4<M>(g, base, nb, L, ll, lane, live, stage, va, vb, vc, vd);
#pragma unroll
    for (int s = 0; s < M; ++s)
        e[s] = Eq<double>{j0 + s > 0 ? va[s] : 0.0, vb[s], j0 + s < n - 1 ? vc[s] : 0.0, vd[s]};
}

// ---- residual check (refine mode with a tolerance) -------------------------
// A system stops taking corrections once max|r| <= tol max|d| over its rows
// (r = d - A x in FP64, the residual of the current x). The check compares
// FP32 bounds: rb >= max|r| (each |fl(r)| scaled by 1 + 2^-23, an upper
// bound for normal r; a NaN counts as infinite, so it fails) and
// dl <= max|d| (rounded toward zero), so a passing system meets the target
// within the rounding of one FP64 product. Checked from the second pass on
// (the first has x = 0).

// The lane's rb from the FP32 right-hand sides of its rows (rows past the
// system hold 0).
template <int M>
__device__ __forceinline__ float rhs_bound(const Eq<float> (&e)[M])
{
    float m = 0.0f;
#pragma unroll
    for (int s = 0; s < M; ++s) {
        const float v = fabsf(e[s].d);
        m = fmaxf(m, v != v ? __builtin_huge_valf() : v);
    }
    // m + m 2^-23 >= m + ulp(m) >= |r| (|fl(r) - r| <= 2^-24 |r|), and
    // m + ulp(m) is representable, so the rounded product stays above |r|
    return m * (1.0f + 0x1p-23f);
}
// The lane's dl (rows past the system or a dead lane: 0).
template <int M>
__device__ __forceinline__ float d_bound(const double (&vd)[M], int j0, int n, bool live)
{
    double m = 0.0;
#pragma unroll
    for (int s = 0; s < M; ++s)
        if (live && j0 + s < n) m = fmax(m, fabs(vd[s]));
    return __double2float_rz(m);
}
// v of the lane CTRL's DPP row shift names (lanes without a source keep v).
template <int CTRL>
__device__ __forceinline__ float dpp_keep(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v),
                                                      CTRL, 0xf, 0xf, false));
}
// Largest v of the wave, uniform: prefix maxima within rows of 16 lanes by
// DPP row shifts, then the four row maxima.
__device__ __forceinline__ float wave_max(float v)
{
    v = fmaxf(v, dpp_keep<0x111>(v));  // row_shr:1
    v = fmaxf(v, dpp_keep<0x112>(v));  // row_shr:2
    v = fmaxf(v, dpp_keep<0x114>(v));  // row_shr:4
    v = fmaxf(v, dpp_keep<0x118>(v));  // row_shr:8
    const int b = __float_as_int(v);
    return fmaxf(fmaxf(__int_as_float(__builtin_amdgcn_readlane(b, 15)),
                       __int_as_float(__builtin_amdgcn_readlane(b, 31))),
                 fmaxf(__int_as_float(__builtin_amdgcn_readlane(b, 47)),
                       __int_as_float(__builtin_amdgcn_readlane(b, 63))));
}
// Largest v over the aligned segment of L lanes (L a power of two <= 64).
__device__ __forceinline__ float seg_max(float v, int L)
{
    for (int k = 1; k < L; k <<= 1) v = fmaxf(v, __shfl_xor(v, k));
    return v;
}
// Wave items: the check of pass `step` for the systems of the wave (L lanes
// each); returns whether every system of the wave is done.
__device__ __forceinline__ bool wave_check(float rb, float& dl, int step, int L, double tol,
                                           bool& done)
{
    rb = seg_max(rb, L);
    if (step == 1) dl = seg_max(dl, L);
    done = done || (double)rb <= tol * (double)dl;
    return __all(done);
}
// Workgroup items: the check of pass `step` (every thread of the workgroup
// calls it). Wave v publishes its bounds in chk[v] (rb) and chk[16 + v]
// (dl, on the first check); after one barrier lane k < waves reads wave k's,
// the groups of W lanes take their maxima (one system each), and a ballot
// gives every lane both its own group's result and whether all are done.
__device__ __forceinline__ bool wg_check(float rb, float dl, int step, int wave, int waves,
                                         int W, int grp, double tol, float* chk, bool& done)
{
    const int lane = threadIdx.x & (kWave - 1);
    rb = wave_max(rb);
    if (step == 1) dl = wave_max(dl);
    if (lane == 0) {
        chk[wave] = rb;
        if (step == 1) chk[16 + wave] = dl;
    }
    wg_barrier<true>();
    const bool mine = lane < waves;
    float r = mine ? chk[lane] : 0.0f;
    float d = mine ? chk[16 + lane] : 0.0f;
    r = seg_max(r, W);
    d = seg_max(d, W);
    const bool ok = !mine || (double)r <= tol * (double)d;
    const uint64_t bal = __ballot(ok);
    done = done || ((bal >> (grp * W)) & 1);
    return bal == ~0ull;
}

// ---- resident refine mode ---------------------------------------------------
// Column element q of lane ll (unknown q*L + ll of the segment, as
// fetch_cols loads it) into the row layout of a home array.
template <int M>
__device__ __forceinline__ void home_put_cols(char* arr, const double (&t)[M], int L, int ll,
                                              int lane)
{
    const int seg_row = lane - ll;
#pragma unroll
    for (int q = 0; q < M; ++q) {
        const int j = q * L + ll;
        *reinterpret_cast<double*>(arr + Home<M>::at(seg_row + j / M, j % M)) = t[q];
    }
}
template <int M>
__device__ __forceinline__ void home_get_row(const char* arr, int lane, double (&v)[M])
{
    if constexpr (M == 1) {
        v[0] = *reinterpret_cast<const double*>(arr + Home<M>::at(lane, 0));
    } else {
#pragma unroll
        for (int k = 0; k < M; k += 2) {
            const double2 w = *reinterpret_cast<const double2*>(arr + Home<M>::at(lane, k));
            v[k] = w.x;
            v[k + 1] = w.y;
        }
    }
}

// One coalesced read of the lane's rows (arguments as ragged_rows): b and d
// end in registers, a and c in the wave's home area (identity rows past n).
template <int M>
__device__ __forceinline__ void gather_resident(const RaggedArgs& g, int64_t base, int nb, int L,
                                                int ll, int lane, bool live, char* home,
                                                double (&vb)[M], double (&vd)[M])
{
    asm volatile("" : "+v"(ll), "+v"(nb));
    double ta[M], tb[M], tc[M], td[M];
    fetch_cols<M>(g.d, base, nb, L, ll, live, 0.0, td);
    fetch_cols<M>(g.b, base, nb, L, ll, live, 1.0, tb);
    fetch_cols<M>(g.a, base, nb, L, ll, live, 0.0, ta);
    fetch_cols<M>(g.c, base, nb, L, ll, live, 0.0, tc);
    char* ha = home;
    char* hc = home + Home<M>::kArrBytes;
    // b and d pass through a's area on their way to the registers
    lds_poison_wave(ha, Home<M>::kArrBytes, lane);
    home_put_cols<M>(ha, tb, L, ll, lane);
    __builtin_amdgcn_wave_barrier();
    home_get_row<M>(ha, lane, vb);
    __builtin_amdgcn_wave_barrier();
    lds_poison_wave(ha, Home<M>::kArrBytes, lane);
    home_put_cols<M>(ha, td, L, ll, lane);
    __builtin_amdgcn_wave_barrier();
    home_get_row<M>(ha, lane, vd);
    __builtin_amdgcn_wave_barrier();
    lds_poison_wave(home, Home<M>::kWaveBytes, lane);
    home_put_cols<M>(ha, ta, L, ll, lane);
    home_put_cols<M>(hc, tc, L, ll, lane);
    __builtin_amdgcn_wave_barrier();
}

// FP32 operator + FP64 residual rows of the lane's unknowns [j0, j0+M) of a
// system of n unknowns, from the resident operator (the residual's operations
// in the order of ragged_rows: d, then b, a, c). Identity rows past n;
// a[first] and c[last] ignored.
template <int M>
__device__ __forceinline__ void resident_rows(const char* home, int lane, const double (&vb)[M],
                                              const double (&vd)[M], const double (&x)[M],
                                              double x_left, double x_right, int j0, int n,
                                              bool live, Eq<float> (&e)[M])
{
    const char* ha = home;
    const char* hc = home + Home<M>::kArrBytes;
#pragma unroll
    for (int k = 0; k < M; k += 2) {
        double av[2], cv[2];
        if constexpr (M == 1) {
            av[0] = *reinterpret_cast<const double*>(ha + Home<M>::at(lane, 0));
            cv[0] = *reinterpret_cast<const double*>(hc + Home<M>::at(lane, 0));
        } else {
            const double2 a2 = *reinterpret_cast<const double2*>(ha + Home<M>::at(lane, k));
            const double2 c2 = *reinterpret_cast<const double2*>(hc + Home<M>::at(lane, k));
            av[0] = a2.x;
            av[1] = a2.y;
            cv[0] = c2.x;
            cv[1] = c2.y;
        }
#pragma unroll
        for (int h = 0; h < 2 && k + h < M; ++h) {
            const int s = k + h;
            const double al = j0 + s > 0 ? av[h] : 0.0;
            const double cr = j0 + s < n - 1 ? cv[h] : 0.0;
            const double xl = s > 0 ? x[s > 0 ? s - 1 : 0] : x_left;
            const double xr = s < M - 1 ? x[s < M - 1 ? s + 1 : 0] : x_right;
            double r = vd[s];
            r = fma(-vb[s], x[s], r);
            r = fma(-al, xl, r);
            r = fma(-cr, xr, r);
            e[s] = Eq<float>{(float)al, (float)vb[s], (float)cr,
                             (live && j0 + s < n) ? (float)r : 0.0f};
        }
    }
}

// The lane's M unknowns back to unknowns [ll*M, ll*M+M) of the segment,
// transposed through the home area's a rows (dead after the last pass),
// written with coalesced stores (nothing past nb).
template <int M>
__device__ __forceinline__ void scatter_resident(double* __restrict__ X, int64_t base, int nb,
                                                 int L, int ll, int lane, bool live, char* home,
                                                 const double (&v)[M])
{
    char* ha = home;
    // addresses formed here, not kept from the gather across the passes
    asm volatile("" : "+v"(ll), "+v"(nb), "+v"(lane));
    const int seg_row = lane - ll;
    __builtin_amdgcn_wave_barrier();
    lds_poison_wave(ha, Home<M>::kArrBytes, lane);
    if constexpr (M == 1) {
        *reinterpret_cast<double*>(ha + Home<M>::at(lane, 0)) = v[0];
    } else {
#pragma unroll
        for (int k = 0; k < M; k += 2)
            *reinterpret_cast<double2*>(ha + Home<M>::at(lane, k)) = double2{v[k], v[k + 1]};
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int q = 0; q < M; ++q) {
        const int j = q * L + ll;
        const double t = *reinterpret_cast<const double*>(ha + Home<M>::at(seg_row + j / M, j % M));
        if (live && j < nb) X[base + j] = t;
    }
    __builtin_amdgcn_wave_barrier();
}

template <int M, bool TOL>
__device__ void wave_item_resident(const RaggedArgs& g, const RaggedItem& it,
                                   const RaggedBucket& bk, int lane, char* home)
{
    asm volatile("" : "+v"(lane));  // formed per item (see wide_item)
    const int L = bk.lanes;
    const int ll = lane & (L - 1);
    const int seg = lane / L;
    const bool live = seg < it.count;
    int64_t base = 0;
    int n = 2;
    if (live) {
        const RaggedSys sy = g.sys[it.first + seg];
        base = sy.base;
        n = (int)sy.n;
    }
    const int j0 = ll * M;
    double vb[M], vd[M];
    gather_resident<M>(g, base, n, L, ll, lane, live, home, vb, vd);
    double x[M];
#pragma unroll
    for (int s = 0; s < M; ++s) x[s] = 0.0;
    Eq<float> e[M];
    float xf[M];
    float dl = 0.0f;
    bool done = false;
    for (int step = 0; step <= g.iters; ++step) {
        const double xl = lane_step_or_zero<-1>(x[M - 1], ll >= 1);
        const double xr = lane_step_or_zero<1>(x[0], ll + 1 < L);
        resident_rows<M>(home, lane, vb, vd, x, xl, xr, j0, n, live, e);
        if constexpr (TOL) {
            if (step > 0) {
                if (step == 1) dl = d_bound<M>(vd, j0, n, live);
                if (wave_check(rhs_bound<M>(e), dl, step, L, g.tol, done)) break;  // uniform
            }
        }
        wave_core<float, M, 1>(e, xf, lane, ll, L, bk.rest);
        if (!done) {
#pragma unroll
            for (int s = 0; s < M; ++s) x[s] += (double)xf[s];
        }
    }
    scatter_resident<M>(g.x, base, n, L, ll, lane, live, home, x);
}

// Every wave of the workgroup calls this (wg_core has barriers). M unknowns
// per lane: 8 (items of 1024..4096 unknowns), 16 (8192).
template <int M, bool TOL>
__device__ void wg_item_resident(const RaggedArgs& g, const RaggedItem& it,
                                 const RaggedBucket& bk, int wave, int waves, int lane,
                                 char* xchg_base, char* home, float* chk)
{
    asm volatile("" : "+v"(lane));  // formed per item (see wide_item)
    // uniform in the workgroup: held in scalar registers, as are the group and
    // the residual check's decisions derived from it
    const int W = __builtin_amdgcn_readfirstlane(bk.lanes / kWave);
    const int grp = wave / W;
    const int w = wave % W;
    const size_t grp_bytes = ragged_group_bytes<float, M>(W);
    char* xchg = xchg_base + grp * grp_bytes;
    double* edge = reinterpret_cast<double*>(xchg + wg_core_lds_bytes<float, M>(W));
    const bool live = grp < it.count;
    int64_t base = 0;
    int n = 2;
    if (live) {
        const RaggedSys sy = g.sys[it.first + grp];
        base = sy.base;
        n = (int)sy.n;
    }
    // this wave's block of the system: unknowns [w*64*M, (w+1)*64*M)
    const int wb = w * kWave * M;
    const int j0 = wb + lane * M;
    double vb[M], vd[M];
    gather_resident<M>(g, base + wb, n - wb, kWave, lane, lane, live, home, vb, vd);
    double x[M];
#pragma unroll
    for (int s = 0; s < M; ++s) x[s] = 0.0;
    Eq<float> e[M];
    float xf[M];
    bool done = false;
    for (int step = 0; step <= g.iters; ++step) {
        // lane-derived values (masks, permute addresses) are cheap: left
        // visible, the compiler keeps them all across the passes and spills
        int ln = lane, jj = j0, nn = n;
        asm volatile("" : "+v"(ln), "+v"(jj), "+v"(nn));
        double xl = 0.0, xr = 0.0;
        if (step > 0) {
            if (ln == 0) edge[2 * w] = x[0];
            if (ln == kWave - 1) edge[2 * w + 1] = x[M - 1];
            wg_barrier<true>();
            const double xl0 = lane_read(x[M - 1], (ln - 1) & (kWave - 1));
            const double xr0 = lane_read(x[0], (ln + 1) & (kWave - 1));
            xl = ln > 0 ? xl0 : (w > 0 ? edge[2 * w - 1] : 0.0);
            xr = ln < kWave - 1 ? xr0 : (w + 1 < W ? edge[2 * w + 2] : 0.0);
        }
        resident_rows<M>(home, ln, vb, vd, x, xl, xr, jj, nn, live, e);
        if constexpr (TOL) {
            if (step > 0) {
                const float dl = step == 1 ? d_bound<M>(vd, jj, nn, live) : 0.0f;
                if (wg_check(rhs_bound<M>(e), dl, step, wave, waves, W, grp, g.tol, chk, done))
                    break;  // every system of the item
            }
        }
        wg_core<float, M, true>(e, xf, w, W, ln, bk.rest, xchg);
        if (!done) {
#pragma unroll
            for (int s = 0; s < M; ++s) x[s] += (double)xf[s];
        }
    }
    // the lane index formed again here, not kept across the passes
    int ls = (int)threadIdx.x & (kWave - 1);
    asm volatile("" : "+v"(ls));
    scatter_resident<M>(g.x, base + wb, n - wb, kWave, ls, ls, live, home, x);
    wg_barrier<true>();  // exchange areas are reused by the next item
}

// ---- wide items (refine mode, 8193..16384 unknowns) ---------------------------
// One system per 1024-thread workgroup, 16 unknowns per lane, x in
// registers; each pass reads the system's FP64 rows again (coalesced, one
// array at a time through the wave's transposition area).
// FP32 operator + FP64 residual rows (as kern_refine.hip's build_rows) of
// the lane's unknowns [j0, j0+M) of a system with n unknowns. `base`, `nb`
// and `j0b` locate the lane's segment for the gathers: the segment's first
// unknown in memory, the unknowns it holds, and the lane's first unknown
// relative to it. Identity rows past n; a[first] and c[last] ignored.
template <int M>
__device__ __forceinline__ void ragged_rows(const RaggedArgs& g, const double (&x)[M],
                                            double x_left, double x_right, int64_t base,
                                            int nb, int j0, int n, int L, int ll,
                                            int lane, bool live, char* stage, Eq<float> (&e)[M],
                                            double& rmax)
{
    double r[M], v[M];
    auto use_b = [&]() {
#pragma unroll
        for (int s = 0; s < M; ++s) {
            r[s] = fma(-v[s], x[s], r[s]);
            e[s].b = (float)v[s];
        }
    };
    auto use_a = [&]() {
#pragma unroll
        for (int s = 0; s < M; ++s) {
            const double al = j0 + s > 0 ? v[s] : 0.0;
            const double xl = s > 0 ? x[s > 0 ? s - 1 : 0] : x_left;
            r[s] = fma(-al, xl, r[s]);
            e[s].a = (float)al;
        }
    };
    auto use_c = [&]() {
#pragma unroll
        for (int s = 0; s < M; ++s) {
            const double cr = j0 + s < n - 1 ? v[s] : 0.0;
            const double xr = s < M - 1 ? x[s < M - 1 ? s + 1 : 0] : x_right;
            r[s] = fma(-cr, xr, r[s]);
            e[s].c = (float)cr;
            const bool row = live && j0 + s < n;
            e[s].d = row ? (float)r[s] : 0.0f;
            if (row) rmax = fmax(rmax, fabs(r[s]));
        }
    };
    // see gather4 for why
    asm volatile("" : "+v"(ll), "+v"(nb));
    if constexpr (M <= 8 && TRID_RAGGED_AHEAD == 2) {
        // two arrays in flight: the next array's loads are issued before
        // the current one is transposed
        double t0[M], t1[M];
        fetch_cols<M>(g.d, base, nb, L, ll, live, 0.0, t0);
        fetch_cols<M>(g.b, base, nb, L, ll, live, 1.0, t1);
        transpose_rows<M>(t0, L, ll, lane, stage, r);
        fetch_cols<M>(g.a, base, nb, L, ll, live, 0.0, t0);
        transpose_rows<M>(t1, L, ll, lane, stage, v);
        use_b();
        fetch_cols<M>(g.c, base, nb, L, ll, live, 0.0, t1);
        transpose_rows<M>(t0, L, ll, lane, stage, v);
        use_a();
        transpose_rows<M>(t1, L, ll, lane, stage, v);
        use_c();
    } else if constexpr (M <= 8) {
        // all loads first (one memory round trip), then the transposes,
        // each array consumed as soon as it is in rows
        double ta[M], tb[M], tc[M], td[M];
        fetch_cols<M>(g.d, base, nb, L, ll, live, 0.0, td);
        fetch_cols<M>(g.b, base, nb, L, ll, live, 1.0, tb);
        fetch_cols<M>(g.a, base, nb, L, ll, live, 0.0, ta);
        fetch_cols<M>(g.c, base, nb, L, ll, live, 0.0, tc);
        transpose_rows<M>(td, L, ll, lane, stage, r);
        transpose_rows<M>(tb, L, ll, lane, stage, v);
        use_b();
        transpose_rows<M>(ta, L, ll, lane, stage, v);
        use_a();
        transpose_rows<M>(tc, L, ll, lane, stage, v);
        use_c();
    } else {
        double t[M];
        fetch_cols<M>(g.d, base, nb, L, ll, live, 0.0, t);
        transpose_rows<M>(t, L, ll, lane, stage, r);
        fetch_cols<M>(g.b, base, nb, L, ll, live, 1.0, t);
        transpose_rows<M>(t, L, ll, lane, stage, v);
        use_b();
        fetch_cols<M>(g.a, base, nb, L, ll, live, 0.0, t);
        transpose_rows<M>(t, L, ll, lane, stage, v);
        use_a();
        fetch_cols<M>(g.c, base, nb, L, ll, live, 0.0, t);
        transpose_rows<M>(t, L, ll, lane, stage, v);
        use_c();
    }
}


template <bool TOL>
__device__ void wide_item(const RaggedArgs& g, const RaggedItem& it, const RaggedBucket& bk,
                          int wave, int lane, char* xchg_base, char* stage, float* chk)
{
    // Lane-derived addresses are cheap to form; left visible, the compiler
    // forms every one of them once for the whole persistent loop and spills
    // them. Made opaque here, they are formed per item.
    asm volatile("" : "+v"(lane));
    constexpr int M = 16;
    const int W = bk.lanes / kWave;  // 16: one system per item
    const int w = wave;
    const size_t grp_bytes = ragged_group_bytes<float, M>(W);
    (void)grp_bytes;
    char* xchg = xchg_base;
    double* edge = reinterpret_cast<double*>(xchg + wg_core_lds_bytes<float, M>(W));
    const bool live = it.count > 0;
    const RaggedSys sy = g.sys[it.first];
    const int64_t base = sy.base;
    const int n = (int)sy.n;
    // this wave's block of the system: unknowns [w*64*M, (w+1)*64*M)
    const int wb = w * kWave * M;
    const int j0 = wb + lane * M;
    double x[M];
#pragma unroll
    for (int s = 0; s < M; ++s) x[s] = 0.0;
    Eq<float> e[M];
    float xf[M];
    float dl = 0.0f;
    bool done = false;
    for (int step = 0; step <= g.iters; ++step) {
        int ln = lane;
        asm volatile("" : "+v"(ln));
        double xl = 0.0, xr = 0.0;
        if (step > 0) {
            if (ln == 0) edge[2 * w] = x[0];
            if (ln == kWave - 1) edge[2 * w + 1] = x[M - 1];
            wg_barrier<true>();
            const double xl0 = lane_read(x[M - 1], (ln - 1) & (kWave - 1));
            const double xr0 = lane_read(x[0], (ln + 1) & (kWave - 1));
            xl = ln > 0 ? xl0 : (w > 0 ? edge[2 * w - 1] : 0.0);
            xr = ln < kWave - 1 ? xr0 : (w + 1 < W ? edge[2 * w + 2] : 0.0);
        }
        double dmax = 0.0;
        ragged_rows<M>(g, x, xl, xr, base + wb, n - wb, j0, n, kWave, ln, ln, live, stage, e,
                       dmax);
        if constexpr (TOL) {
            if (step == 0) dl = __double2float_rz(dmax);  // x = 0: the rows' |r| are |d|
            else if (wg_check(rhs_bound<M>(e), dl, step, w, W, W, 0, g.tol, chk, done)) break;
        }
        wg_core<float, M, true>(e, xf, w, W, ln, bk.rest, xchg);
        if (!done) {
#pragma unroll
            for (int s = 0; s < M; ++s) x[s] += (double)xf[s];
        }
    }
    scatter_rows<M>(g.x, base + wb, n - wb, kWave, lane, lane, live, stage, x);
    wg_barrier<true>();  // exchange areas are reused by the next item
}

// ---- FP64 mode items (one solve) -------------------------------------------------
template <int M>
__device__ void wave_item_direct(const RaggedArgs& g, const RaggedItem& it,
                                 const RaggedBucket& bk, int lane, char* stage)
{
    asm volatile("" : "+v"(lane));  // formed per item (see wide_item)
    const int L = bk.lanes;
    const int ll = lane & (L - 1);
    const int seg = lane / L;
    const bool live = seg < it.count;
    int64_t base = 0;
    int n = 2;
    if (live) {
        const RaggedSys sy = g.sys[it.first + seg];
        base = sy.base;
        n = (int)sy.n;
    }
    const int j0 = ll * M;
    double x[M];
    Eq<double> e[M];
    direct_rows<M>(g, base, n, j0, n, L, ll, lane, live, stage, e);
    wave_core<double, M, 1>(e, x, lane, ll, L, bk.rest);
    scatter_rows<M>(g.x, base, n, L, ll, lane, live, stage, x);
}

// Every wave of the workgroup calls this (wg_core has barriers); M as in
// wg_item_resident.
template <int M>
__device__ void wg_item_direct(const RaggedArgs& g, const RaggedItem& it, const RaggedBucket& bk,
                               int wave, int lane, char* xchg_base, char* stage)
{
    asm volatile("" : "+v"(lane));  // formed per item (see wide_item)
    const int W = bk.lanes / kWave;
    const int grp = wave / W;
    const int w = wave % W;
    const size_t grp_bytes = ragged_group_bytes<double, M>(W);
    char* xchg = xchg_base + grp * grp_bytes;
    const bool live = grp < it.count;
    int64_t base = 0;
    int n = 2;
    if (live) {
        const RaggedSys sy = g.sys[it.first + grp];
        base = sy.base;
        n = (int)sy.n;
    }
    // this wave's block of the system: unknowns [w*64*M, (w+1)*64*M)
    const int wb = w * kWave * M;
    const int j0 = wb + lane * M;
    double x[M];
    Eq<double> e[M];
    direct_rows<M>(g, base + wb, n - wb, j0, n, kWave, lane, lane, live, stage, e);
    wg_core<double, M, true>(e, x, w, W, lane, bk.rest, xchg);
    scatter_rows<M>(g.x, base + wb, n - wb, kWave, lane, lane, live, stage, x);
    wg_barrier<true>();  // exchange areas are reused by the next item
}

// Items [lo, hi) from the atomic queue counter `ctr`; every thread of the
// workgroup calls it, `body(item)` ends with a workgroup barrier. The next
// item's queue slot is claimed when an item starts and published when it
// ends, so the atomic's latency hides behind the item. The barriers of the
// workgroup items order LDS only (wg_barrier<true>): an item's x stores
// (rows no other item touches) and the next slot's atomic stay in flight
// across them instead of being waited for at every barrier, as
// __syncthreads() would (its release fence waits for all memory operations).
template <typename F>
__device__ __forceinline__ void drain_queue(int lo, int hi, int* ctr, int* s_item, F&& body)
{
    if (threadIdx.x == 0) s_item[0] = lo + atomicAdd(ctr, 1);
    __syncthreads();
    int item = s_item[0];
    for (int k = 1; item < hi; ++k) {
        int next = 0;
        if (threadIdx.x == 0) next = lo + atomicAdd(ctr, 1);
        body(item);
        // body ends with a barrier, after which every thread has read
        // s_item[(k - 1) & 1]; slot k & 1 was last read before that barrier
        if (threadIdx.x == 0) s_item[k & 1] = next;
        wg_barrier<true>();
        item = s_item[k & 1];
    }
}

}  // namespace

__global__ void ragged_reset_kernel(int* counters)
{
    if (threadIdx.x < 3) counters[threadIdx.x] = 0;
}

namespace {

// Workgroup items [lo, hi) of a WAVES-wave workgroup at M unknowns per lane,
// claimed from `ctr`.
template <int MODE, int WAVES, int M>
__device__ __forceinline__ void wg_items(const RaggedArgs& g, int lo, int hi, int* ctr)
{
    __shared__ __align__(16) char smem[wg_lds_bytes<MODE, WAVES, M>()];
    __shared__ float s_chk[32];  // residual check: per-wave bounds (refine mode)
    __shared__ int s_item[2];
    __shared__ RaggedBucket s_bk[kRaggedMaxBuckets];
    // the wave index is uniform in the wave: held in a scalar register, so
    // that the group, the system's base and size derived from it are too
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int lane = threadIdx.x & (kWave - 1);
    char* area = smem + (size_t)wave * ragged_wave_bytes<MODE, M>();
    char* xchg_base = smem + (size_t)WAVES * ragged_wave_bytes<MODE, M>();
    lds_poison_wg(smem, (unsigned)sizeof(smem));
    lds_poison_wg(s_bk, (unsigned)sizeof(s_bk));
    if (threadIdx.x < g.n_buckets) s_bk[threadIdx.x] = g.buckets[threadIdx.x];
    drain_queue(lo, hi, ctr, s_item, [&](int item) {
        const RaggedItem it = g.items[item];
        const RaggedBucket bk = s_bk[it.bucket];
        if constexpr (MODE != kDirect)
            wg_item_resident<M, MODE == kRefineTol>(g, it, bk, wave, WAVES, lane, xchg_base,
                                                    area, s_chk);
        else wg_item_direct<M>(g, it, bk, wave, lane, xchg_base, area);
    });
}

}  // namespace

// Items at 16 unknowns per lane [n_wide_items, n_full_items): 8-wave
// workgroups (256 registers; one per CU) that hold one 8192-unknown system
// or 8/W systems of W waves (buckets the table gives 16 per lane). On 4,096
// uniform systems of 8,192 unknowns this shape solves in 124 us FP32 /
// 338 us FP64 against 190 / 357 us for 16 waves of 8 per lane.
template <int MODE>
__global__ __launch_bounds__(kFullThreads) void ragged_wg_kernel(RaggedArgs g)
{
    wg_items<MODE, kFullWaves, 16>(g, g.n_wide_items, g.n_full_items, &g.counters[1]);
}

// Workgroup items of at most 8 waves [n_full_items, n_wg_items), largest
// bucket first; at most 128 registers per lane, so that two workgroups share
// a CU.
template <int MODE>
__global__ __launch_bounds__(kHalfThreads) __attribute__((amdgpu_waves_per_eu(4)))
void ragged_half_kernel(RaggedArgs g)
{
    wg_items<MODE, kHalfWaves, 8>(g, g.n_full_items, g.n_wg_items, &g.counters[2]);
}

// Wide items [0, n_wide_items) (refine mode; TOL: with the residual check).
template <bool TOL>
__global__ __launch_bounds__(kWgThreads) void ragged_wide_kernel(RaggedArgs g)
{
    __shared__ __align__(16) char smem[kWideLdsBytes];
    __shared__ float s_chk[32];  // residual check: per-wave bounds
    __shared__ int s_item[2];
    __shared__ RaggedBucket s_bk[kRaggedMaxBuckets];
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));  // uniform
    const int lane = threadIdx.x & (kWave - 1);
    lds_poison_wg(smem, (unsigned)sizeof(smem));
    lds_poison_wg(s_bk, (unsigned)sizeof(s_bk));
    if (threadIdx.x < g.n_buckets) s_bk[threadIdx.x] = g.buckets[threadIdx.x];
    drain_queue(0, g.n_wide_items, &g.counters[0], s_item, [&](int item) {
        const RaggedItem it = g.items[item];
        wide_item<TOL>(g, it, s_bk[it.bucket], wave, lane, smem + kWideStageBytes,
                  smem + (size_t)wave * RStage<16>::kBytes, s_chk);
    });
}

// Wave items, each wave on its own, dealt round-robin over every wave of the
// grid (largest bucket first, so each wave gets a similar mix). A shared
// queue counter here would be claimed by thousands of waves at once, and each
// claim would queue behind the others at one L2 address: on 15,000 items of
// 512 unknowns that cost 5x the solve.
template <int MODE>
__global__ __launch_bounds__(kWaveThreads) void ragged_wave_kernel(RaggedArgs g)
{
    __shared__ __align__(16) char smem[(size_t)kWaveWaves * ragged_wave_bytes<MODE, 8>()];
    __shared__ RaggedBucket s_bk[kRaggedMaxBuckets];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    char* area = smem + (size_t)wave * ragged_wave_bytes<MODE, 8>();
    lds_poison_wg(smem, (unsigned)sizeof(smem));
    lds_poison_wg(s_bk, (unsigned)sizeof(s_bk));
    if (threadIdx.x < g.n_buckets) s_bk[threadIdx.x] = g.buckets[threadIdx.x];
    __syncthreads();
    const int waves_total = (int)gridDim.x * kWaveWaves;
    for (int witem = (int)blockIdx.x * kWaveWaves + wave; witem < g.n_wave_items;
         witem += waves_total) {
        const RaggedItem it = g.items[g.n_wg_items + witem];
        const RaggedBucket bk = s_bk[it.bucket];
        if constexpr (MODE != kDirect) {
            constexpr bool TOL = MODE == kRefineTol;
            switch (bk.per_lane) {
                case 1: wave_item_resident<1, TOL>(g, it, bk, lane, area); break;
                case 2: wave_item_resident<2, TOL>(g, it, bk, lane, area); break;
                case 4: wave_item_resident<4, TOL>(g, it, bk, lane, area); break;
                case 8: wave_item_resident<8, TOL>(g, it, bk, lane, area); break;
                default: break;
            }
        } else {
            switch (bk.per_lane) {
                case 1: wave_item_direct<1>(g, it, bk, lane, area); break;
                case 2: wave_item_direct<2>(g, it, bk, lane, area); break;
                case 4: wave_item_direct<4>(g, it, bk, lane, area); break;
                case 8: wave_item_direct<8>(g, it, bk, lane, area); break;
                default: break;
            }
        }
    }
}

namespace {

int slots_of(const void* kern, int threads, int device)
{
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, 0) != hipSuccess)
        per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        cus = 256;
    if (per_cu < 1) per_cu = 1;
    return per_cu * cus;
}

int clamp_grid(int64_t work, int slots)
{
    return (int)(work < slots ? (work > 0 ? work : 1) : slots);
}

}  // namespace

RaggedGrids ragged_grids(int device, const RaggedArgs& g)
{
    // workgroup slots of each kernel in each mode (refine, FP64, refine with
    // the residual check) on each device (the current device when first
    // asked), found once: [mode][wide, 8192, half, wave]
    static std::mutex mu;
    static int cache[64][3][4];
    static bool known[64];
    const int dv = device >= 0 && device < 64 ? device : 0;
    const int mode = g.direct ? kDirect : (g.tol > 0.0 ? kRefineTol : kRefine);
    int slots[4];
    {
        std::lock_guard<std::mutex> lk(mu);
        int (*c)[4] = cache[dv];
        if (!known[dv]) {
            auto fill = [&](int m, const void* wg, const void* half, const void* wave) {
                c[m][1] = slots_of(wg, kFullThreads, device);
                c[m][2] = slots_of(half, kHalfThreads, device);
                c[m][3] = slots_of(wave, kWaveThreads, device);
            };
            fill(kRefine, (const void*)ragged_wg_kernel<kRefine>,
                 (const void*)ragged_half_kernel<kRefine>, (const void*)ragged_wave_kernel<kRefine>);
            fill(kDirect, (const void*)ragged_wg_kernel<kDirect>,
                 (const void*)ragged_half_kernel<kDirect>, (const void*)ragged_wave_kernel<kDirect>);
            fill(kRefineTol, (const void*)ragged_wg_kernel<kRefineTol>,
                 (const void*)ragged_half_kernel<kRefineTol>,
                 (const void*)ragged_wave_kernel<kRefineTol>);
            c[kRefine][0] = slots_of((const void*)ragged_wide_kernel<false>, kWgThreads, device);
            c[kDirect][0] = 1;  // no FP64 form of the wide items
            c[kRefineTol][0] = slots_of((const void*)ragged_wide_kernel<true>, kWgThreads, device);
            known[dv] = true;
        }
        for (int k = 0; k < 4; ++k) slots[k] = c[mode][k];
    }
    RaggedGrids r{};
    r.wide = clamp_grid(g.n_wide_items, slots[0]);
    r.wg = clamp_grid(g.n_full_items - g.n_wide_items, slots[1]);
    r.half = clamp_grid(g.n_wg_items - g.n_full_items, slots[2]);
    r.wave = clamp_grid((g.n_wave_items + kWaveWaves - 1) / kWaveWaves, slots[3]);
    return r;
}

template <int MODE>
static hipError_t launch_ragged_mode(const RaggedArgs& g, const RaggedGrids& grid,
                                     hipStream_t stream)
{
    hipError_t e = hipSuccess;
    if constexpr (MODE != kDirect) {
        if (g.n_wide_items > 0) {
            hipLaunchKernelGGL((ragged_wide_kernel<MODE == kRefineTol>), dim3((unsigned)grid.wide),
                               dim3(kWgThreads), 0, stream, g);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
    }
    if (g.n_full_items > g.n_wide_items) {
        hipLaunchKernelGGL(ragged_wg_kernel<MODE>, dim3((unsigned)grid.wg), dim3(kFullThreads), 0,
                           stream, g);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (g.n_wg_items > g.n_full_items) {
        hipLaunchKernelGGL(ragged_half_kernel<MODE>, dim3((unsigned)grid.half),
                           dim3(kHalfThreads), 0, stream, g);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (g.n_wave_items > 0) {
        hipLaunchKernelGGL(ragged_wave_kernel<MODE>, dim3((unsigned)grid.wave),
                           dim3(kWaveThreads), 0, stream, g);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_ragged(const RaggedArgs& g, const RaggedGrids& grid, hipStream_t stream)
{
    if (g.direct && g.n_wide_items > 0) return hipErrorInvalidValue;  // no 16384 FP64 form
    if (g.direct && g.tol != 0.0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(ragged_reset_kernel, dim3(1), dim3(64), 0, stream, g.counters);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (g.direct) return launch_ragged_mode<kDirect>(g, grid, stream);
    if (g.tol > 0.0) return launch_ragged_mode<kRefineTol>(g, grid, stream);
    return launch_ragged_mode<kRefine>(g, grid, stream);
}

}  // namespace trid
