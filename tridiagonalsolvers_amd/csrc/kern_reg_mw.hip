// kern_reg_mw.hip -- register-resident CR / CR-PCR for systems longer than one
// wave can hold (n = W * 64 * M, W = 2..16 waves per workgroup).
//
// Methods: TRID_CR, TRID_CRPCR (cr_kernel.cu, crpcr_kernel.cu) for n up to
// 16 * 64 * 16 = 16384 (FP32) / 8192 (FP64). One workgroup owns one system;
// wave w owns unknowns [w*64*M, (w+1)*64*M), lane l of it M consecutive ones.
//
// The CR forward levels only ever reach to the RIGHT across a wave boundary:
// a reduced equation i = k*2s + 2s - 1 reads i - s and i + s, and i - s never
// leaves the block that i tops. So every wave reduces its own block in
// lockstep; only the top equation of each wave (lane 63, slot M-1) reads the
// next wave, and its chain of values is provisional until that wave's
// "prefix" equations are known. Each wave publishes its prefix equations
// (lane 0's slot 2^(k-1)-1 after in-lane level k-1, lane 2^(j-1)-1's top
// after cross-lane level j-1) to LDS; after one barrier every wave recomputes
// its lane-63 chain from its saved level-0 top equation. The reduced system
// (one equation per `rest`-th position) is then solved:
//   rest <= W: one wave gathers the W wave-top equations and runs CR down to
//              `rest` plus PCR, cross-lane, with no barrier;
//   rest  > W: PCR across the workgroup through LDS (two barriers a level),
//              then 2x2 Cramer solves.
// Back-substitution runs cross-lane, then in-lane; the only value a wave
// needs from another is the previous wave's top unknown.
#include "reg_core.hpp"
#include "reg_io.hpp"
#include "trid_device.hpp"
#include "trid_launch.hpp"

namespace trid {

template <typename T, int M, int IO>
__device__ __forceinline__ void reg_mw_body(const T* __restrict__ A, const T* __restrict__ B,
                                            const T* __restrict__ C, const T* __restrict__ D,
                                            T* __restrict__ X, int64_t n, int rest, int vec,
                                            int64_t N, int64_t bs)
{
    constexpr int E = 16 / (int)sizeof(T);
    const int W = blockDim.x / kWave;
    const int w = threadIdx.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t sys = blockIdx.x;
    const int64_t base = sys * bs;  // bs = n for a contiguous batch
    const int64_t total = (N - 1) * bs + n;
    const int64_t j0 = (int64_t)(w * kWave + lane) * M;

    // LDS: the staging rows (load and store) share space with the exchange
    // area (prefix equations, reduced system, solved tops) used in between.
    extern __shared__ __align__(16) char trid_smem[];
    char* lds_stage = trid_smem + w * Stage<T, M>::kWaveBytes;

    // ---- load -------------------------------------------------------------
    Eq<T> e[M];
    bool loaded = false;
    if constexpr (M % E == 0 && IO >= 2) {
        if (vec) {
            load_staged<T, M, IO == 3>(A, B, C, D, base + (int64_t)w * kWave * M, total, lane,
                                       lds_stage, e, kNoConstCoef<T>);
            loaded = true;
        }
    } else if constexpr (M % E == 0) {
        if (vec) {
            load_direct<T, M, IO == 0>(A, B, C, D, base + j0, e, kNoConstCoef<T>);
            loaded = true;
        }
    }
    if (!loaded) {
#pragma unroll
        for (int s = 0; s < M; ++s) {
            const int64_t j = j0 + s;
            if (j < n)
                e[s] = Eq<T>{A[base + j], B[base + j], C[base + j], D[base + j]};
            else e[s] = identity_eq<T>();
        }
    }
    if (w == 0 && lane == 0) e[0].a = T(0);
#pragma unroll
    for (int s = 0; s < M; ++s)
        if (j0 + s == n - 1) e[s].c = T(0);
    T xv[M];
    wg_core<T, M>(e, xv, w, W, lane, rest, trid_smem);

    // ---- store ----------------------------------------------------------------
    __syncthreads();  // exchange area becomes staging rows again
    if constexpr (M % E == 0 && IO >= 2) {
        if (vec) {
            // the store addresses are formed here from the thread index
            // (cheap); kept from before the solve they would stay live
            // across it, and at 16 unknowns per lane the core needs every
            // register
            int tid = (int)threadIdx.x;
            asm volatile("" : "+v"(tid));
            const int w2 = tid / kWave;
            char* stage2 = trid_smem + w2 * Stage<T, M>::kWaveBytes;
            store_staged<T, M, IO == 3>(X, (int64_t)blockIdx.x * bs + (int64_t)w2 * kWave * M,
                                        (N - 1) * bs + n, tid & (kWave - 1), stage2, xv);
            return;
        }
    } else if constexpr (M % E == 0) {
        if (vec) {
            store_direct<T, M, IO == 0>(X, base + j0, xv);
            return;
        }
    }
#pragma unroll
    for (int s = 0; s < M; ++s) {
        const int64_t j = j0 + s;
        if (j < n) X[base + j] = xv[s];
    }
}

// Up to 16 waves per system (the register budget of 1024 threads: 128 per
// lane) ...
template <typename T, int M, int IO>
__global__ __launch_bounds__(1024) void reg_mw_kernel(const T* __restrict__ A,
                                                      const T* __restrict__ B,
                                                      const T* __restrict__ C,
                                                      const T* __restrict__ D,
                                                      T* __restrict__ X, int64_t n, int rest,
                                                      int vec, int64_t N, int64_t bs)
{
    reg_mw_body<T, M, IO>(A, B, C, D, X, n, rest, vec, N, bs);
}

// ... or up to 8 waves with twice the registers: FP64 at 16 unknowns per lane.
template <typename T, int M, int IO>
__global__ __launch_bounds__(512) void reg_mw512_kernel(const T* __restrict__ A,
                                                        const T* __restrict__ B,
                                                        const T* __restrict__ C,
                                                        const T* __restrict__ D,
                                                        T* __restrict__ X, int64_t n, int rest,
                                                        int vec, int64_t N, int64_t bs)
{
    reg_mw_body<T, M, IO>(A, B, C, D, X, n, rest, vec, N, bs);
}

// ---- persistent form with the next system's a, c copied ahead -------------
//
// Systems of exactly kPfWaves * 64 * 16 = 8192 unknowns at 16 per lane: in
// FP64 one system fills a CU's registers, in FP32 two do, so in the kernels
// above the CU's memory pipe idles while its systems are solved. Here each
// workgroup walks the systems blockIdx.x, blockIdx.x + gridDim.x, ... and,
// while it solves one, copies the next one's a and c (2 x 64 KiB FP64,
// 2 x 32 KiB FP32) from HBM straight into LDS (global_load_lds_dwordx4:
// 16 bytes a lane, no registers). b and d are loaded after the wait,
// coalesced, and transposed through the same LDS area (free by then); x is
// transposed through the exchange area and stored coalesced, as many waves
// at a time as that area holds.
//
// Copy layout: copy instruction k of wave w moves the k-th KiB of the wave's
// block (64 lanes x 16 elements) in lane order, so the LDS holds the block
// as it is in HBM and lane l reads its row at l * 16 * sizeof(T). Those
// 16-byte reads meet bank conflicts, but they overlap the b, d loads.
// Measured on 4,096 x 8,192 (kernel average): FP64 275-279 us against 334 us
// for reg_mw512_kernel, FP32 111 us against 119 us for reg_mw_kernel (the
// time of a plain stream of the same bytes); FP64 with the rows swizzled
// within each KiB (conflict-free reads, 16-byte pieces out of order within
// each copy instruction) 281 us; x by per-lane 16-byte stores (lanes 128
// bytes apart) 551-565 us.
constexpr int kPfWaves = 8;
constexpr int kPfM = 16;
constexpr int kPfN = kPfWaves * kWave * kPfM;  // 8192 unknowns
template <typename T>
constexpr size_t kPfArrayBytes = (size_t)kPfN * sizeof(T);  // 64 KiB FP64 / 32 KiB FP32
template <typename T>
constexpr size_t kPfWaveBytes = (size_t)kWave * kPfM * sizeof(T);  // 8 / 4 KiB

// Issue the copies of system `sys`'s a and c (wave w's blocks) into `pf`:
// a at pf, c at pf + kPfArrayBytes, wave w's block at w * kPfWaveBytes.
template <typename T>
__device__ __forceinline__ void pf_copy(const T* __restrict__ A, const T* __restrict__ C,
                                        int64_t sys, int w, int lane, char* pf)
{
    constexpr int E = 16 / (int)sizeof(T);                  // elements a lane copies
    constexpr int K = (int)(kPfWaveBytes<T> / 1024);        // KiB of a wave's block
    const int64_t wbase = sys * kPfN + (int64_t)w * kWave * kPfM;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int64_t el = wbase + (int64_t)kWave * E * k + E * lane;
        char* dst_a = pf + (size_t)w * kPfWaveBytes<T> + (size_t)k * 1024;
        char* dst_c = dst_a + kPfArrayBytes<T>;
        __builtin_amdgcn_global_load_lds(
            (const void*)(A + el), (__attribute__((address_space(3))) void*)dst_a, 16, 0, 0);
        __builtin_amdgcn_global_load_lds(
            (const void*)(C + el), (__attribute__((address_space(3))) void*)dst_c, 16, 0, 0);
    }
}

// LDS of the persistent kernel: the exchange area, and the total
template <typename T>
constexpr size_t pf_xchg_bytes() { return wg_core_lds_bytes<T, kPfM>(kPfWaves); }
template <typename T>
constexpr size_t pf_lds_bytes() { return 2 * kPfArrayBytes<T> + pf_xchg_bytes<T>(); }
// workgroups per CU: FP64 fills the registers (256 a lane); FP32 takes two
// (128 registers a lane, twice 81,152 bytes of LDS)
template <typename T>
constexpr int kPfPerCu = sizeof(T) == 8 ? 1 : 2;

template <typename T>
__global__ __launch_bounds__(kPfWaves * kWave)
__attribute__((amdgpu_waves_per_eu(2 * kPfPerCu<T>))) void reg_mw_pf_kernel(
    const T* __restrict__ A, const T* __restrict__ B, const T* __restrict__ C,
    const T* __restrict__ D, T* __restrict__ X, int rest, int64_t N)
{
    constexpr int M = kPfM, W = kPfWaves;
    using S = Stage<T, M>;
    static_assert(2 * kPfArrayBytes<T> >= (size_t)W * S::kWaveBytes,
                  "the b, d staging rows fit in the copy area");
    static_assert(kPfPerCu<T> * pf_lds_bytes<T>() <= 160 * 1024, "LDS of the CU's workgroups");
    // Two static LDS variables: the compiler gives each its own alias scope,
    // so it can tell the exchange area's accesses from the copies' target and
    // does not make them wait for the copies in flight.
    __shared__ __align__(16) char pf[2 * kPfArrayBytes<T>];  // a, c copies; b, d staging
    __shared__ __align__(16) char xchg[pf_xchg_bytes<T>()];  // the core's exchange area
    // the wave index is uniform: a scalar register, and so are the wave's
    // base offsets derived from it
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int lane = threadIdx.x & (kWave - 1);

    int64_t sys = blockIdx.x;
    pf_copy<T>(A, C, sys, w, lane, pf);
    for (; sys < N; sys += gridDim.x) {
        const int64_t wbase = sys * kPfN + (int64_t)w * kWave * M;
        // this system's copies have landed (every wave's): s_waitcnt vmcnt(0)
        // by the builtin, which the compiler's own wait insertion sees (so it
        // adds no wait for the copies before the reads below)
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        wg_barrier<true>();
        // a, c rows from the copies
        T ta[M], tb[M], tc[M], td[M];
        {
            const char* ra =
                pf + (size_t)w * kPfWaveBytes<T> + (size_t)lane * (kPfM * sizeof(T));
#pragma unroll
            for (int q = 0; q < S::Q; ++q) {
                const u32x4 va = *reinterpret_cast<const u32x4*>(ra + 16 * q);
                const u32x4 vc = *reinterpret_cast<const u32x4*>(ra + kPfArrayBytes<T> + 16 * q);
                __builtin_memcpy(&ta[q * S::E], &va, 16);
                __builtin_memcpy(&tc[q * S::E], &vc, 16);
            }
        }
        // b, d: coalesced 16-byte loads (instruction q: the q-th KiB)
        u32x4 rb[S::Q], rd[S::Q];
#pragma unroll
        for (int q = 0; q < S::Q; ++q) {
            const int64_t el = wbase + (int64_t)(kWave * q + lane) * S::E;
            rb[q] = gload16<true>(B + el);
            rd[q] = gload16<true>(D + el);
        }
        wg_barrier<true>();  // the copy area is free: staging rows for b, d
        {
            char* st = pf + (size_t)w * S::kWaveBytes;
            lds_poison_wave(st, S::kWaveBytes, lane);
#pragma unroll
            for (int q = 0; q < S::Q; ++q)
                *reinterpret_cast<u32x4*>(st + S::chunk_addr(kWave * q + lane)) = rb[q];
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int q = 0; q < S::Q; ++q) {
                const u32x4 t = *reinterpret_cast<const u32x4*>(st + lane * S::kRowBytes + q * 16);
                __builtin_memcpy(&tb[q * S::E], &t, 16);
            }
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int q = 0; q < S::Q; ++q)
                *reinterpret_cast<u32x4*>(st + S::chunk_addr(kWave * q + lane)) = rd[q];
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int q = 0; q < S::Q; ++q) {
                const u32x4 t = *reinterpret_cast<const u32x4*>(st + lane * S::kRowBytes + q * 16);
                __builtin_memcpy(&td[q * S::E], &t, 16);
            }
        }
        wg_barrier<true>();  // staging done: the next copies may land
        if (sys + gridDim.x < N) pf_copy<T>(A, C, sys + gridDim.x, w, lane, pf);

        Eq<T> e[M];
#pragma unroll
        for (int s = 0; s < M; ++s) e[s] = Eq<T>{ta[s], tb[s], tc[s], td[s]};
        if (w == 0 && lane == 0) e[0].a = T(0);
        if (w == W - 1 && lane == kWave - 1) e[M - 1].c = T(0);
        T xv[M];
        wg_core<T, M, true>(e, xv, w, W, lane, rest, xchg);
        // x through the exchange area, as many waves at a time as it holds
        constexpr int kPer = (int)(pf_xchg_bytes<T>() / S::kWaveBytes);
        static_assert(kPer >= 1, "one wave's staging rows fit the exchange area");
        wg_barrier<true>();  // every wave is past its last exchange-area read
        for (int r = 0; r < W; r += kPer) {
            if (w >= r && w < r + kPer)
                store_staged<T, M, true>(X, wbase, N * kPfN, lane,
                                         xchg + (size_t)(w - r) * S::kWaveBytes, xv);
            if (r + kPer < W) wg_barrier<true>();
        }
    }
}

// The persistent form serves dense batches of 8192-unknown systems with
// 16-byte aligned arrays, at 16 unknowns per lane.
template <typename T>
static bool pf_form(const RegArgs<T>& g)
{
    return g.n == kPfN && g.M == kPfM && g.L == kPfWaves * kWave && g.vec &&
           (g.bs == 0 || g.bs == g.n) && g.io >= 2;
}

template <typename T>
static hipError_t launch_pf(const RegArgs<T>& g)
{
    const int64_t slots = (int64_t)kPfPerCu<T> * device_cus();
    const int grid = (int)(g.N < slots ? g.N : slots);
    hipLaunchKernelGGL(reg_mw_pf_kernel<T>, dim3((unsigned)grid), dim3(kPfWaves * kWave), 0,
                       g.stream, g.a, g.b, g.c, g.d, g.x, g.rest, g.N);
    return hipGetLastError();
}

template <typename T, int M>
size_t mw_lds_bytes(int W, int io, int vec)
{
    const size_t xchg = wg_core_lds_bytes<T, M>(W);
    const size_t stage = (io >= 2 && vec) ? (size_t)W * Stage<T, M>::kWaveBytes : 0;
    return xchg > stage ? xchg : stage;
}

template <typename T, int M, int IO>
static hipError_t launch_mw_io(const RegArgs<T>& g)
{
    if constexpr (M == kPfM && IO >= 2) {
        if (pf_form(g)) return launch_pf(g);
    }
    const int W = g.L / kWave;
    const size_t lds = mw_lds_bytes<T, M>(W, IO, g.vec);
    if (lds > 160 * 1024) return hipErrorInvalidConfiguration;
    // FP64 at 16 unknowns per lane needs the 512-thread form (W <= 8)
    constexpr bool wide = sizeof(T) == 8 && M == 16;
    if (wide && W > 8) return hipErrorInvalidValue;
    const void* kern;
    if constexpr (wide) kern = (const void*)reg_mw512_kernel<T, M, IO>;
    else kern = (const void*)reg_mw_kernel<T, M, IO>;
    if (lds > 64 * 1024) {
        hipError_t e =
            hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    if constexpr (wide)
        hipLaunchKernelGGL((reg_mw512_kernel<T, M, IO>), dim3((unsigned)g.N), dim3(kWave * W),
                           lds, g.stream, g.a, g.b, g.c, g.d, g.x, g.n, g.rest, g.vec, g.N,
                           g.bs ? g.bs : g.n);
    else
        hipLaunchKernelGGL((reg_mw_kernel<T, M, IO>), dim3((unsigned)g.N), dim3(kWave * W), lds,
                           g.stream, g.a, g.b, g.c, g.d, g.x, g.n, g.rest, g.vec, g.N,
                           g.bs ? g.bs : g.n);
    return hipGetLastError();
}

template <typename T, int M>
static hipError_t launch_mw_m(const RegArgs<T>& g)
{
    switch (g.io) {
        case 0: return launch_mw_io<T, M, 0>(g);
        case 1: return launch_mw_io<T, M, 1>(g);
        case 2: return launch_mw_io<T, M, 2>(g);
        case 3: return launch_mw_io<T, M, 3>(g);
        default: return hipErrorInvalidValue;
    }
}

// g.L = lanes per system = 64 * W (W = 2..16 waves), g.M unknowns per lane.
template <typename T>
hipError_t launch_reg_mw(const RegArgs<T>& g)
{
    if (g.L % kWave != 0 || g.L / kWave < 2 || g.L / kWave > 16) return hipErrorInvalidValue;
    if (g.N <= 0 || g.N > 0x7fffffffLL) return hipErrorInvalidConfiguration;
    switch (g.M) {
        case 2: return launch_mw_m<T, 2>(g);
        case 4: return launch_mw_m<T, 4>(g);
        case 8: return launch_mw_m<T, 8>(g);
        case 16: return launch_mw_m<T, 16>(g);
        default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_reg_mw<float>(const RegArgs<float>&);
template hipError_t launch_reg_mw<double>(const RegArgs<double>&);

}  // namespace trid
