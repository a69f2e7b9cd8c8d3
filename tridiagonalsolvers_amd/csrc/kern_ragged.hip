// kern_ragged.hip -- one persistent launch over a ragged batch.
//
// No counterpart in the reference (every reference wrapper takes one
// systemSize for the whole batch). A ragged batch is numSystems systems of
// arbitrary sizes n_s >= 2 in flat FP64 arrays, system s occupying
// [offsets[s], offsets[s+1]). Every system is solved in FP32 on the register
// CR-PCR core and refined in FP64 (the kern_refine.hip scheme), padded to
// the power of two above n_s with identity rows (exact).
//
// Work queue (built on the host, trid_ragged.cpp): systems are bucketed by
// padded size; each bucket has a configuration (unknowns per lane M, lanes
// per system, CR->PCR switch size) chosen by the autotune table.
//   * wave items: buckets up to 512 unknowns; one item = up to 64/L systems
//     of one bucket, solved by one wave, L lanes each (no barrier);
//   * workgroup items: buckets of 1024..16384 unknowns; 8 unknowns per lane
//     (16 for 16384), W = n/512 waves per system (16 for 16384), 16/W systems
//     per item, one per group of W waves of the 1024-thread workgroup
//     (barriers are workgroup-wide, and every group of an item runs the same
//     bucket's code, so they match).
// The grid is persistent (one workgroup per CU slot). Workgroups first drain
// the workgroup items, largest bucket first, from an atomic queue counter
// that a small kernel resets before every launch (so the pair can be
// captured in a hipGraph and replayed); then each wave works through the
// wave items dealt to it round-robin.
#include "reg_core.hpp"
#include "trid_device.hpp"
#include "trid_launch.hpp"

#include <type_traits>

namespace trid {

namespace {

// Coefficient arrays whose loads are in flight at once in the refinement
// passes (2 or 4). Measured on the 100,000-system batch: sequential 3.43 ms,
// 4 in flight 3.68 ms (the registers spill), 2 in flight 3.34 ms. The FP64
// mode always issues all four (1.95 ms, against 2.09 ms sequential).
#ifndef TRID_RAGGED_AHEAD
#define TRID_RAGGED_AHEAD 2
#endif

constexpr int kRaggedThreads = 1024;
constexpr int kRaggedWaves = kRaggedThreads / kWave;
static_assert(kRaggedWaves == 16, "workgroup items assume 16 waves");

// Transposition area of one wave (coalesced loads and stores, below).
// Row r holds the M consecutive unknowns of lane r, padded by 16 bytes:
// a row stride of M*8+16 bytes puts the 16 lanes of one pass of a 16-byte
// LDS access on disjoint banks. M = 16 is staged in two rounds of 32 rows.
template <int M>
struct RStage {
    static constexpr int kRowBytes = M * 8 + 16;
    static constexpr int kRounds = M > 8 ? M / 8 : 1;
    static constexpr int kQ = M / kRounds;         // loads per lane per round
    static constexpr int kRows = kWave / kRounds;  // rows per round
    static constexpr int kBytes = kRows * kRowBytes;
};
constexpr int kStageWaveBytes = RStage<8>::kBytes;  // 5120, the largest
static_assert(RStage<16>::kBytes <= kStageWaveBytes, "stage area");

// Kernel modes: FP32 solve + FP64 refinement (workgroup items at 8
// unknowns per lane, or also 16 in the wide form for 16384-unknown
// systems), or one FP64 solve (8 unknowns per lane, systems up to 8192).
enum RaggedMode { kRefine = 0, kRefineWide = 1, kDirect = 2 };

// LDS of the workgroup items: 16/W groups of W waves, each with the
// exchange area of wg_core and 2 W boundary unknowns (wg_core_lds_bytes is
// linear in W, so the 16-wave group is the largest). After it, one
// transposition area per wave.
template <int MODE>
constexpr size_t ragged_xchg_bytes()
{
    return MODE == kDirect ? wg_core_lds_bytes<double, 8>(16) + 2 * 16 * sizeof(double)
                           : wg_core_lds_bytes<float, 16>(16) + 2 * 16 * sizeof(double);
}
template <int MODE>
constexpr size_t ragged_lds_bytes()
{
    return ragged_xchg_bytes<MODE>() + (size_t)kRaggedWaves * kStageWaveBytes;
}
static_assert(ragged_lds_bytes<kDirect>() <= 160 * 1024, "LDS of the FP64 mode");
static_assert(ragged_lds_bytes<kRefineWide>() <= 160 * 1024, "LDS of the refine mode");

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

// FP32 operator + FP64 residual rows (as kern_refine.hip's build_rows) of
// the lane's unknowns [j0, j0+M) of a system with n unknowns. `base`, `nb`
// and `j0b` locate the lane's segment for the gathers: the segment's first
// unknown in memory, the unknowns it holds, and the lane's first unknown
// relative to it. Identity rows past n; a[first] and c[last] ignored.
template <int M>
__device__ __forceinline__ void ragged_rows(const RaggedArgs& g, const double (&x)[M],
                                            double x_left, double x_right, int64_t base,
                                            int nb, int j0, int n, int L, int ll,
                                            int lane, bool live, char* stage, Eq<float> (&e)[M])
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
            e[s].d = (live && j0 + s < n) ? (float)r[s] : 0.0f;
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

template <bool DIRECT, int M>
__device__ void wave_item(const RaggedArgs& g, const RaggedItem& it, const RaggedBucket& bk,
                          int lane, char* stage)
{
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
    if constexpr (DIRECT) {
        Eq<double> e[M];
        direct_rows<M>(g, base, n, j0, n, L, ll, lane, live, stage, e);
        wave_core<double, M, 1>(e, x, lane, ll, L, bk.rest);
        scatter_rows<M>(g.x, base, n, L, ll, lane, live, stage, x);
        return;
    }
#pragma unroll
    for (int s = 0; s < M; ++s) x[s] = 0.0;
    Eq<float> e[M];
    float xf[M];
    for (int step = 0; step <= g.iters; ++step) {
        const double xl = lane_step_or_zero<-1>(x[M - 1], ll >= 1);
        const double xr = lane_step_or_zero<1>(x[0], ll + 1 < L);
        ragged_rows<M>(g, x, xl, xr, base, n, j0, n, L, ll, lane, live, stage, e);
        wave_core<float, M, 1>(e, xf, lane, ll, L, bk.rest);
#pragma unroll
        for (int s = 0; s < M; ++s) x[s] += (double)xf[s];
    }
    scatter_rows<M>(g.x, base, n, L, ll, lane, live, stage, x);
}

// All 16 waves of the workgroup call this (wg_core has barriers).
template <bool DIRECT, int M>
__device__ void wg_item(const RaggedArgs& g, const RaggedItem& it, const RaggedBucket& bk,
                        int wave, int lane, char* smem, char* stage)
{
    using TC = typename std::conditional<DIRECT, double, float>::type;  // core type
    const int W = bk.lanes / kWave;
    const int grp = wave / W;
    const int w = wave % W;
    const size_t grp_bytes = wg_core_lds_bytes<TC, M>(W) + (size_t)2 * W * sizeof(double);
    char* xchg = smem + grp * grp_bytes;
    double* edge = reinterpret_cast<double*>(xchg + wg_core_lds_bytes<TC, M>(W));
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
    if constexpr (DIRECT) {
        Eq<double> e[M];
        direct_rows<M>(g, base + wb, n - wb, j0, n, kWave, lane, lane, live, stage, e);
        wg_core<double, M>(e, x, w, W, lane, bk.rest, xchg);
        scatter_rows<M>(g.x, base + wb, n - wb, kWave, lane, lane, live, stage, x);
        __syncthreads();  // exchange areas are reused by the next item
        return;
    }
#pragma unroll
    for (int s = 0; s < M; ++s) x[s] = 0.0;
    Eq<float> e[M];
    float xf[M];
    for (int step = 0; step <= g.iters; ++step) {
        double xl = 0.0, xr = 0.0;
        if (step > 0) {
            if (lane == 0) edge[2 * w] = x[0];
            if (lane == kWave - 1) edge[2 * w + 1] = x[M - 1];
            __syncthreads();
            const double xl0 = lane_read(x[M - 1], (lane - 1) & (kWave - 1));
            const double xr0 = lane_read(x[0], (lane + 1) & (kWave - 1));
            xl = lane > 0 ? xl0 : (w > 0 ? edge[2 * w - 1] : 0.0);
            xr = lane < kWave - 1 ? xr0 : (w + 1 < W ? edge[2 * w + 2] : 0.0);
        }
        ragged_rows<M>(g, x, xl, xr, base + wb, n - wb, j0, n, kWave, lane, lane, live, stage,
                       e);
        wg_core<float, M>(e, xf, w, W, lane, bk.rest, xchg);
#pragma unroll
        for (int s = 0; s < M; ++s) x[s] += (double)xf[s];
    }
    scatter_rows<M>(g.x, base + wb, n - wb, kWave, lane, lane, live, stage, x);
    __syncthreads();  // exchange areas are reused by the next item
}

}  // namespace

__global__ void ragged_reset_kernel(int* counters)
{
    if (threadIdx.x < 2) counters[threadIdx.x] = 0;
}

// Modes (RaggedMode): the wide refine form handles 16384-unknown systems
// (16 unknowns per lane in the workgroup items); it needs more registers
// than the others, so it is a separate instantiation and the common one
// keeps its register budget.
template <int MODE>
__global__ __launch_bounds__(kRaggedThreads) void ragged_kernel(RaggedArgs g)
{
    constexpr bool DIRECT = MODE == kDirect;
    __shared__ __align__(16) char smem[ragged_lds_bytes<MODE>()];
    __shared__ int s_item[2];
    __shared__ RaggedBucket s_bk[kRaggedMaxBuckets];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    char* stage = smem + ragged_xchg_bytes<MODE>() + (size_t)wave * kStageWaveBytes;
    lds_poison_wg(smem, (unsigned)sizeof(smem));
    lds_poison_wg(s_bk, (unsigned)sizeof(s_bk));
    if (threadIdx.x < g.n_buckets) s_bk[threadIdx.x] = g.buckets[threadIdx.x];

    // phase 1: workgroup items (long systems), largest bucket first. The
    // next item's queue slot is claimed when an item starts and published
    // when it ends, so the atomic's latency hides behind the item.
    if (threadIdx.x == 0) s_item[0] = atomicAdd(&g.counters[0], 1);
    __syncthreads();
    int item = s_item[0];
    for (int k = 1; item < g.n_wg_items; ++k) {
        int next = 0;
        if (threadIdx.x == 0) next = atomicAdd(&g.counters[0], 1);
        const RaggedItem it = g.items[item];
        const RaggedBucket bk = s_bk[it.bucket];
        bool done = false;
        if constexpr (MODE == kRefineWide) {
            if (bk.per_lane == 16) {
                wg_item<false, 16>(g, it, bk, wave, lane, smem, stage);
                done = true;
            }
        }
        if (!done) wg_item<DIRECT, 8>(g, it, bk, wave, lane, smem, stage);
        // wg_item ends with a barrier, after which every thread has read
        // s_item[(k - 1) & 1]; slot k & 1 was last read before that barrier
        if (threadIdx.x == 0) s_item[k & 1] = next;
        __syncthreads();
        item = s_item[k & 1];
    }
    // phase 2: wave items, each wave on its own, dealt round-robin over every
    // wave of the grid (largest bucket first, so each wave gets a similar
    // mix). A shared queue counter here would be claimed by thousands of
    // waves at once, and each claim would queue behind the others at one L2
    // address: on 15,000 items of 512 unknowns that cost 5x the solve.
    const int waves_total = (int)gridDim.x * kRaggedWaves;
    for (int witem = (int)blockIdx.x * kRaggedWaves + wave; witem < g.n_wave_items;
         witem += waves_total) {
        const RaggedItem it = g.items[g.n_wg_items + witem];
        const RaggedBucket bk = s_bk[it.bucket];
        switch (bk.per_lane) {
            case 1: wave_item<DIRECT, 1>(g, it, bk, lane, stage); break;
            case 2: wave_item<DIRECT, 2>(g, it, bk, lane, stage); break;
            case 4: wave_item<DIRECT, 4>(g, it, bk, lane, stage); break;
            case 8: wave_item<DIRECT, 8>(g, it, bk, lane, stage); break;
            default: break;
        }
    }
}

hipError_t launch_ragged(const RaggedArgs& g, int grid, hipStream_t stream)
{
    hipLaunchKernelGGL(ragged_reset_kernel, dim3(1), dim3(64), 0, stream, g.counters);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (g.direct) {
        if (g.wide) return hipErrorInvalidValue;  // no 16384-unknown FP64 form
        hipLaunchKernelGGL(ragged_kernel<kDirect>, dim3((unsigned)grid), dim3(kRaggedThreads), 0,
                           stream, g);
    } else if (g.wide) {
        hipLaunchKernelGGL(ragged_kernel<kRefineWide>, dim3((unsigned)grid),
                           dim3(kRaggedThreads), 0, stream, g);
    } else {
        hipLaunchKernelGGL(ragged_kernel<kRefine>, dim3((unsigned)grid), dim3(kRaggedThreads), 0,
                           stream, g);
    }
    return hipGetLastError();
}

// Persistent grid: every workgroup slot of the device.
int ragged_max_grid(int device)
{
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ragged_kernel<kRefine>,
                                                     kRaggedThreads, 0) !=
        hipSuccess)
        per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        cus = 256;
    if (per_cu < 1) per_cu = 1;
    return per_cu * cus;
}

}  // namespace trid
