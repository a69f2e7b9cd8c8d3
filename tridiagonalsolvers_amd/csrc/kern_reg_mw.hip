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
