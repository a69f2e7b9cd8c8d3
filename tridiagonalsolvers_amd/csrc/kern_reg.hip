// kern_reg.hip -- register-resident CR / CR-PCR / PCR for gfx950.
//
// Methods: TRID_CR, TRID_CRPCR, TRID_PCR (cr_kernel.cu, crpcr_kernel.cu,
// pcr_kernel.cu in the reference). The reference keeps a whole system in
// shared memory, one thread per two unknowns, and pays a barrier per level.
// Here a system is owned by a segment of L lanes of one wave64, each lane
// holding M consecutive unknowns in registers:
//
//   1. in-lane CR levels      (stride 2..M; the top equation of a lane reads
//                              the right lane's "prefix" equation by bpermute)
//   2. cross-lane CR levels   (while more than `rest` equations remain)
//   3. PCR on the `rest` system, cross-lane, all reads before writes because
//      every level works on registers
//   4. 2x2 Cramer solves at lane distance L/2
//   5. CR back-substitution across lanes, then inside lanes
//
// No LDS storage and no barrier is used. The arithmetic of each step is the
// reference's (reduce_eq / solve2 / backsub in trid_device.hpp; PCR levels
// past one equation per lane run on rows scaled to a unit diagonal, the same
// elimination divided through: reduce_neq), and the sequence of
// steps is exactly CR-PCR with switch size `rest`:
// rest = 2 is CR (crpcr with no PCR level equals cr_kernel.cu), rest = n is
// PCR (pcr_kernel.cu: log2(n/2) PCR levels, then n/2 2x2 solves).
//
// Loads and stores: each lane moves its M consecutive elements with 16-byte
// vector accesses when the batch is dense (n == L*M, 16-byte aligned);
// otherwise element-wise with identity rows past the end (any n >= 2).
#include "reg_core.hpp"
#include "reg_io.hpp"
#include "trid_device.hpp"
#include "trid_launch.hpp"

namespace trid {


// T: element type. M: unknowns per lane. R: equations per lane handed to PCR
// (R = 1: the PCR system has at most one equation per lane, rest <= L;
//  R > 1: rest = R * L). FULL (R > 1 only): one system per wave, L = 64.
template <typename T, int M, int R, int IO, bool FULL>
__global__ __launch_bounds__(256) void reg_crpcr_kernel(const T* __restrict__ A,
                                                        const T* __restrict__ B,
                                                        const T* __restrict__ C,
                                                        const T* __restrict__ D,
                                                        T* __restrict__ X, int64_t n,
                                                        int64_t N, int L, int logL, int rest,
                                                        int vec, int64_t bs)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int ll = lane & (L - 1);
    const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t sys = gtid >> logL;
    const bool live = sys < N;
    const int64_t base = sys * bs;  // bs = n for a contiguous batch
    const int64_t j0 = (int64_t)ll * M;

    constexpr int E = 16 / (int)sizeof(T);
    extern __shared__ __align__(16) char trid_smem[];
    char* lds = trid_smem + (threadIdx.x / kWave) * Stage<T, M>::kWaveBytes;
    // dense: a wave's systems are one block of 64*M elements (contiguous
    // batches), or 64/L systems of n = L*M elements, bs apart (strided)
    const int64_t wave_base =
        bs == n ? (gtid / kWave) * (int64_t)(kWave * M) : ((gtid / kWave) * kWave >> logL) * bs;
    const int64_t total = (N - 1) * bs + n;
    const int seg_log2 = (bs == n || L == kWave) ? -1 : logL + Log2<M>::value;

    // Batches that are not 16-byte aligned (vec = 0) are launched with IO = 0,
    // the one instantiation that carries the element-wise path: a second
    // load path merged into the others raises their register pressure
    // through the whole solve (FP32 PCR: 256 VGPRs with it, 121 without).
    const bool scalar = M % E != 0 || (IO == 0 && !vec);
    Eq<T> e[M];
    if constexpr (M % E == 0 && IO >= 2) {
        load_staged<T, M, IO == 3>(A, B, C, D, wave_base, total, lane, lds, e, kNoConstCoef<T>,
                                   seg_log2, bs);
        if (!live) set_identity<T, M>(e);
    } else if constexpr (M % E == 0) {
        if (!scalar) {
            if (live) load_direct<T, M, IO == 0>(A, B, C, D, base + j0, e, kNoConstCoef<T>);
            else set_identity<T, M>(e);
        }
    }
    if (scalar) {
#pragma unroll
        for (int s = 0; s < M; ++s) {
            const int64_t j = j0 + s;
            if (live && j < n)
                e[s] = Eq<T>{A[base + j], B[base + j], C[base + j], D[base + j]};
            else e[s] = identity_eq<T>();
        }
    }
    // a[first] and c[last] are ignored (treated as 0), as in the gold
    // Thomas (tridiagonal_gold.h:28).
    if (ll == 0) e[0].a = T(0);
#pragma unroll
    for (int s = 0; s < M; ++s)
        if (j0 + s == n - 1) e[s].c = T(0);

    T xv[M];
    wave_core<T, M, R, FULL>(e, xv, lane, ll, L, rest);

    if constexpr (M % E == 0 && IO >= 2) {
        store_staged<T, M, IO == 3>(X, wave_base, total, lane, lds, xv, seg_log2, bs);
        return;
    } else if constexpr (M % E == 0) {
        if (!scalar) {
            if (live) store_direct<T, M, IO == 0>(X, base + j0, xv);
            return;
        }
    }
    if (live) {
#pragma unroll
        for (int s = 0; s < M; ++s) {
            const int64_t j = j0 + s;
            if (j < n) X[base + j] = xv[s];
        }
    }
}

template <typename T, int M, int IO>
static hipError_t launch_reg_io(const RegArgs<T>& g, int R, dim3 grid, dim3 block)
{
    const int logL = __builtin_ctz((unsigned)g.L);
    const size_t lds = (IO >= 2 && g.vec) ? (size_t)(block.x / kWave) * Stage<T, M>::kWaveBytes : 0;
    const bool full = g.L == kWave;
#define TRID_REG_LAUNCH(RR, FF)                                                                \
    hipLaunchKernelGGL((reg_crpcr_kernel<T, M, RR, IO, FF>), grid, block, lds, g.stream, g.a, \
                       g.b, g.c, g.d, g.x, g.n, g.N, g.L, logL, g.rest, g.vec,               \
                       g.bs ? g.bs : g.n)
#define TRID_REG_CASE(RR)                                                                      \
    case RR:                                                                                   \
        if constexpr (RR <= M) {                                                               \
            if (RR > 1 && full) TRID_REG_LAUNCH(RR, (RR > 1));                                \
            else TRID_REG_LAUNCH(RR, false);                                                   \
            return hipGetLastError();                                                          \
        }                                                                                      \
        break;
    switch (R) {
        TRID_REG_CASE(1)
        TRID_REG_CASE(2)
        TRID_REG_CASE(4)
        TRID_REG_CASE(8)
        TRID_REG_CASE(16)
        default: break;
    }
#undef TRID_REG_CASE
#undef TRID_REG_LAUNCH
    return hipErrorInvalidValue;
}

template <typename T, int M>
static hipError_t launch_reg_m(const RegArgs<T>& g, int R, dim3 grid, dim3 block)
{
    switch (g.vec ? g.io : 0) {  // element-wise batches: IO = 0 (see the kernel)
        case 0: return launch_reg_io<T, M, 0>(g, R, grid, block);
        case 1: return launch_reg_io<T, M, 1>(g, R, grid, block);
        case 2: return launch_reg_io<T, M, 2>(g, R, grid, block);
        case 3: return launch_reg_io<T, M, 3>(g, R, grid, block);
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t launch_reg_crpcr(const RegArgs<T>& g)
{
    const int R = g.rest > g.L ? g.rest / g.L : 1;
    const int64_t threads = g.N * (int64_t)g.L;
    const int block = 256;
    const int64_t blocks = (threads + block - 1) / block;
    if (blocks <= 0 || blocks > 0x7fffffffLL) return hipErrorInvalidConfiguration;
    dim3 grid((unsigned)blocks), blk(block);
    switch (g.M) {
        case 1: return launch_reg_m<T, 1>(g, R, grid, blk);
        case 2: return launch_reg_m<T, 2>(g, R, grid, blk);
        case 4: return launch_reg_m<T, 4>(g, R, grid, blk);
        case 8: return launch_reg_m<T, 8>(g, R, grid, blk);
        case 16: return launch_reg_m<T, 16>(g, R, grid, blk);
        default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_reg_crpcr<float>(const RegArgs<float>&);
template hipError_t launch_reg_crpcr<double>(const RegArgs<double>&);

}  // namespace trid
