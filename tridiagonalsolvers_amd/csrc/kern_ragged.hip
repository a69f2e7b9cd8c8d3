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
// the workgroup items, largest bucket first, then each wave drains wave
// items on its own; both queues are atomic counters that a small kernel
// resets before every launch, so the pair can be captured in a hipGraph and
// replayed.
#include "reg_core.hpp"
#include "trid_device.hpp"
#include "trid_launch.hpp"

namespace trid {

namespace {

constexpr int kRaggedThreads = 1024;
constexpr int kRaggedWaves = kRaggedThreads / kWave;
static_assert(kRaggedWaves == 16, "workgroup items assume 16 waves");
// LDS of the workgroup items: 16/W groups of W waves, each with the
// exchange area of wg_core and 2 W boundary unknowns. The largest case is
// 16 waves at 16 unknowns per lane (16384-unknown systems).
constexpr size_t kRaggedLds = 16 * (10 * 16 + 64 * 16 + 64 * 4 + 16);

// FP32 operator + FP64 residual rows (as kern_refine.hip's build_rows) for a
// ragged system starting at `base` with n unknowns; element-wise loads.
template <int M>
__device__ __forceinline__ void ragged_rows(const double* __restrict__ A,
                                            const double* __restrict__ B,
                                            const double* __restrict__ C,
                                            const double* __restrict__ D, const double (&x)[M],
                                            double x_left, double x_right, int64_t base,
                                            int64_t j0, int64_t n, bool live, Eq<float> (&e)[M])
{
#pragma unroll
    for (int s = 0; s < M; ++s) {
        const int64_t j = j0 + s;
        const bool row = live && j < n;
        const double va = row && j > 0 ? A[base + j] : 0.0;
        const double vb = row ? B[base + j] : 1.0;
        const double vc = row && j < n - 1 ? C[base + j] : 0.0;
        const double vd = row ? D[base + j] : 0.0;
        const double xl = s > 0 ? x[s > 0 ? s - 1 : 0] : x_left;
        const double xr = s < M - 1 ? x[s < M - 1 ? s + 1 : 0] : x_right;
        double r = fma(-vb, x[s], vd);
        r = fma(-va, xl, r);
        r = fma(-vc, xr, r);
        e[s] = Eq<float>{(float)va, (float)vb, (float)vc, row ? (float)r : 0.0f};
    }
}

template <int M>
__device__ void wave_item(const RaggedArgs& g, const RaggedItem& it, const RaggedBucket& bk,
                          int lane)
{
    const int L = bk.lanes;
    const int ll = lane & (L - 1);
    const int seg = lane / L;
    const bool live = seg < it.count;
    int64_t base = 0, n = 2;
    if (live) {
        const int s = g.order[it.first + seg];
        base = g.offsets[s];
        n = g.offsets[s + 1] - base;
    }
    const int64_t j0 = (int64_t)ll * M;
    double x[M];
#pragma unroll
    for (int s = 0; s < M; ++s) x[s] = 0.0;
    Eq<float> e[M];
    float xf[M];
    for (int step = 0; step <= g.iters; ++step) {
        const double xl = lane_read_or_zero(x[M - 1], lane - 1, ll >= 1);
        const double xr = lane_read_or_zero(x[0], lane + 1, ll + 1 < L);
        ragged_rows<M>(g.a, g.b, g.c, g.d, x, xl, xr, base, j0, n, live, e);
        wave_core<float, M, 1>(e, xf, lane, ll, L, bk.rest);
#pragma unroll
        for (int s = 0; s < M; ++s) x[s] += (double)xf[s];
    }
    if (!live) return;
#pragma unroll
    for (int s = 0; s < M; ++s)
        if (j0 + s < n) g.x[base + j0 + s] = x[s];
}

// All 16 waves of the workgroup call this (wg_core has barriers).
template <int M>
__device__ void wg_item(const RaggedArgs& g, const RaggedItem& it, const RaggedBucket& bk,
                        int wave, int lane, char* smem)
{
    const int W = bk.lanes / kWave;
    const int grp = wave / W;
    const int w = wave % W;
    const size_t grp_bytes = wg_core_lds_bytes<float, M>(W) + (size_t)2 * W * sizeof(double);
    char* xchg = smem + grp * grp_bytes;
    double* edge = reinterpret_cast<double*>(xchg + wg_core_lds_bytes<float, M>(W));
    const bool live = grp < it.count;
    int64_t base = 0, n = 2;
    if (live) {
        const int s = g.order[it.first + grp];
        base = g.offsets[s];
        n = g.offsets[s + 1] - base;
    }
    const int64_t j0 = (int64_t)(w * kWave + lane) * M;
    double x[M];
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
        ragged_rows<M>(g.a, g.b, g.c, g.d, x, xl, xr, base, j0, n, live, e);
        wg_core<float, M>(e, xf, w, W, lane, bk.rest, xchg);
#pragma unroll
        for (int s = 0; s < M; ++s) x[s] += (double)xf[s];
    }
    if (live) {
#pragma unroll
        for (int s = 0; s < M; ++s)
            if (j0 + s < n) g.x[base + j0 + s] = x[s];
    }
    __syncthreads();  // exchange areas are reused by the next item
}

}  // namespace

__global__ void ragged_reset_kernel(int* counters)
{
    if (threadIdx.x < 2) counters[threadIdx.x] = 0;
}

// WIDE: the plan has 16384-unknown systems (16 unknowns per lane in the
// workgroup items). That path needs more registers than the others, so it is
// a separate instantiation and the common one keeps its register budget.
template <bool WIDE>
__global__ __launch_bounds__(kRaggedThreads) void ragged_kernel(RaggedArgs g)
{
    __shared__ __align__(16) char smem[kRaggedLds];
    __shared__ int s_item;
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);

    // phase 1: workgroup items (long systems), largest bucket first
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(&g.counters[0], 1);
        __syncthreads();
        const int item = s_item;
        __syncthreads();
        if (item >= g.n_wg_items) break;
        const RaggedItem it = g.items[item];
        const RaggedBucket bk = g.buckets[it.bucket];
        if constexpr (WIDE) {
            if (bk.per_lane == 16) {
                wg_item<16>(g, it, bk, wave, lane, smem);
                continue;
            }
        }
        wg_item<8>(g, it, bk, wave, lane, smem);
    }
    // phase 2: wave items, each wave on its own
    for (;;) {
        int item = 0;
        if (lane == 0) item = atomicAdd(&g.counters[1], 1);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item >= g.n_wave_items) break;
        const RaggedItem it = g.items[g.n_wg_items + item];
        const RaggedBucket bk = g.buckets[it.bucket];
        switch (bk.per_lane) {
            case 1: wave_item<1>(g, it, bk, lane); break;
            case 2: wave_item<2>(g, it, bk, lane); break;
            case 4: wave_item<4>(g, it, bk, lane); break;
            case 8: wave_item<8>(g, it, bk, lane); break;
            default: break;
        }
    }
}

hipError_t launch_ragged(const RaggedArgs& g, int grid, hipStream_t stream)
{
    hipLaunchKernelGGL(ragged_reset_kernel, dim3(1), dim3(64), 0, stream, g.counters);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (g.wide)
        hipLaunchKernelGGL(ragged_kernel<true>, dim3((unsigned)grid), dim3(kRaggedThreads), 0,
                           stream, g);
    else
        hipLaunchKernelGGL(ragged_kernel<false>, dim3((unsigned)grid), dim3(kRaggedThreads), 0,
                           stream, g);
    return hipGetLastError();
}

// Persistent grid: every workgroup slot of the device.
int ragged_max_grid(int device)
{
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ragged_kernel<false>,
                                                     kRaggedThreads, 0) !=
        hipSuccess)
        per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        cus = 256;
    if (per_cu < 1) per_cu = 1;
    return per_cu * cus;
}

}  // namespace trid
