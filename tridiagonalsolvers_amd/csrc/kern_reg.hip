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
// reference's (reduce_eq / solve2 / backsub in trid_device.hpp), and the
// sequence of steps is exactly CR-PCR with switch size `rest`:
// rest = 2 is CR (crpcr with no PCR level equals cr_kernel.cu), rest = n is
// PCR (pcr_kernel.cu: log2(n/2) PCR levels, then n/2 2x2 solves).
//
// Loads and stores: each lane moves its M consecutive elements with 16-byte
// vector accesses when the batch is dense (n == L*M, 16-byte aligned);
// otherwise element-wise with identity rows past the end (any n >= 2).
#include "reg_io.hpp"
#include "trid_device.hpp"
#include "trid_launch.hpp"

namespace trid {


// T: element type. M: unknowns per lane. R: equations per lane handed to PCR
// (R = 1: the PCR system has at most one equation per lane, rest <= L;
//  R > 1: rest = R * L).
template <typename T, int M, int R, int IO>
__global__ __launch_bounds__(256) void reg_crpcr_kernel(const T* __restrict__ A,
                                                        const T* __restrict__ B,
                                                        const T* __restrict__ C,
                                                        const T* __restrict__ D,
                                                        T* __restrict__ X, int64_t n,
                                                        int64_t N, int L, int logL, int rest,
                                                        int vec)
{
    constexpr int S = M / R;                // in-lane stride of the PCR equations
    constexpr int KIN = Log2<S>::value;     // in-lane CR levels
    const int lane = threadIdx.x & (kWave - 1);
    const int ll = lane & (L - 1);
    const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t sys = gtid >> logL;
    const bool live = sys < N;
    const int64_t base = sys * n;
    const int64_t j0 = (int64_t)ll * M;

    constexpr int E = 16 / (int)sizeof(T);
    extern __shared__ __align__(16) char trid_smem[];
    char* lds = trid_smem + (threadIdx.x / kWave) * Stage<T, M>::kWaveBytes;
    const int64_t wave_base = (gtid / kWave) * (int64_t)(kWave * M);  // dense: n == L*M
    const int64_t total = n * N;

    Eq<T> e[M];
    if constexpr (M % E == 0 && IO >= 2) {
        if (vec) {
            load_staged<T, M, IO == 3>(A, B, C, D, wave_base, total, lane, lds, e);
            if (!live) set_identity<T, M>(e);
        }
    } else if constexpr (M % E == 0) {
        if (vec) {
            if (live) load_direct<T, M, IO == 0>(A, B, C, D, base + j0, e);
            else set_identity<T, M>(e);
        }
    }
    if (!vec || M % E != 0) {
#pragma unroll
        for (int s = 0; s < M; ++s) {
            const int64_t j = j0 + s;
            if (live && j < n) e[s] = Eq<T>{A[base + j], B[base + j], C[base + j], D[base + j]};
            else e[s] = identity_eq<T>();
        }
    }
    // a[first] and c[last] are ignored (treated as 0), as in the gold
    // Thomas (tridiagonal_gold.h:28).
    if (ll == 0) e[0].a = T(0);
#pragma unroll
    for (int s = 0; s < M; ++s)
        if (j0 + s == n - 1) e[s].c = T(0);

    // ---- 1. in-lane CR forward levels -----------------------------------
#pragma unroll
    for (int k = 1; k <= KIN; ++k) {
        const int SS = 1 << k, H = SS >> 1;
        // The top equation of this lane needs the right lane's equation at
        // slot H-1, which that lane finished at level k-1.
        const Eq<T> nb = lane_read_eq(e[H - 1], lane + 1, ll + 1 < L);
#pragma unroll
        for (int j = 0; j < M / SS; ++j) {
            const int i = SS * j + SS - 1;
            const int ir = (i + H < M) ? i + H : i;
            const Eq<T> rgt = (i + H < M) ? e[ir] : nb;
            e[i] = reduce_eq(e[i], e[i - H], rgt);
        }
    }

    T xv[M];
#pragma unroll
    for (int s = 0; s < M; ++s) xv[s] = T(0);

    if constexpr (R == 1) {
        // ---- 2. cross-lane CR levels on slot M-1 while count > rest ------
        int hs = 1;
        for (int count = L; count > rest; count >>= 1, hs <<= 1) {
            const bool active = ((ll + 1) & (2 * hs - 1)) == 0;
            const Eq<T> lft = lane_read_eq(e[M - 1], lane - hs, ll - hs >= 0);
            const Eq<T> rgt = lane_read_eq(e[M - 1], lane + hs, ll + hs < L);
            const Eq<T> r = reduce_eq(e[M - 1], lft, rgt);
            if (active) e[M - 1] = r;
        }
        // hs = L / rest: spacing of the lanes that hold the rest system
        const int LS = hs;
        const bool pact = ((ll + 1) & (LS - 1)) == 0;
        // ---- 3. PCR: log2(rest/2) levels, lane distance LS .. L/4 ----------
        for (int dist = LS; dist < (L >> 1); dist <<= 1) {
            const Eq<T> lft = lane_read_eq(e[M - 1], lane - dist, ll - dist >= 0);
            const Eq<T> rgt = lane_read_eq(e[M - 1], lane + dist, ll + dist < L);
            const Eq<T> r = reduce_eq(e[M - 1], lft, rgt);
            if (pact) e[M - 1] = r;
        }
        // ---- 4. 2x2 solves at lane distance L/2 --------------------------
        const int half = L >> 1;
        const bool lower = ll < half;
        const Eq<T> partner = lane_read_eq(e[M - 1], lower ? lane + half : lane - half, true);
        {
            const Eq<T> lo = lower ? e[M - 1] : partner;
            const Eq<T> hi = lower ? partner : e[M - 1];
            T x1, x2;
            solve2(lo, hi, x1, x2);
            xv[M - 1] = lower ? x1 : x2;
        }
        // ---- 5a. cross-lane CR back-substitution -------------------------
        for (int h = LS >> 1; h >= 1; h >>= 1) {
            const T xl = lane_read_or_zero(xv[M - 1], lane - h, ll - h >= 0);
            const T xr = lane_read_or_zero(xv[M - 1], lane + h, ll + h < L);
            const bool solve = ((ll + 1) & (2 * h - 1)) == h;
            const T v = backsub(e[M - 1], xl, xr);
            if (solve) xv[M - 1] = v;
        }
    } else {
        // rest = R * L: lane ll holds PCR equations q = ll*R + r at slots
        // S*r + S - 1, r = 0..R-1.
        // ---- 3a. PCR levels with distance < R (partly inside the lane) ---
#pragma unroll
        for (int p = 0; (1 << p) < R; ++p) {
            const int dl = 1 << p;
            Eq<T> fromL[R], fromR[R];
#pragma unroll
            for (int q = 0; q < R; ++q) {
                if (q < dl) {
                    fromL[q] = lane_read_eq(e[S * (R - dl + q) + S - 1], lane - 1, ll >= 1);
                    fromR[q] = lane_read_eq(e[S * q + S - 1], lane + 1, ll + 1 < L);
                }
            }
            Eq<T> nw[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const Eq<T> lft = (r - dl >= 0) ? e[S * ((r - dl >= 0) ? r - dl : 0) + S - 1]
                                                : fromL[(r - dl >= 0) ? 0 : r];
                const Eq<T> rgt = (r + dl < R) ? e[S * ((r + dl < R) ? r + dl : 0) + S - 1]
                                               : fromR[(r + dl < R) ? 0 : r + dl - R];
                nw[r] = reduce_eq(e[S * r + S - 1], lft, rgt);
            }
#pragma unroll
            for (int r = 0; r < R; ++r) e[S * r + S - 1] = nw[r];
        }
        // ---- 3b. PCR levels with distance >= R: lane distance 1 .. L/4 ---
        for (int ld = 1; ld < (L >> 1); ld <<= 1) {
            Eq<T> nw[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const Eq<T> lft = lane_read_eq(e[S * r + S - 1], lane - ld, ll - ld >= 0);
                const Eq<T> rgt = lane_read_eq(e[S * r + S - 1], lane + ld, ll + ld < L);
                nw[r] = reduce_eq(e[S * r + S - 1], lft, rgt);
            }
#pragma unroll
            for (int r = 0; r < R; ++r) e[S * r + S - 1] = nw[r];
        }
        // ---- 4. 2x2 solves at lane distance L/2 --------------------------
        const int half = L >> 1;
        const bool lower = ll < half;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const Eq<T> own = e[S * r + S - 1];
            const Eq<T> partner = lane_read_eq(own, lower ? lane + half : lane - half, true);
            const Eq<T> lo = lower ? own : partner;
            const Eq<T> hi = lower ? partner : own;
            T x1, x2;
            solve2(lo, hi, x1, x2);
            xv[S * r + S - 1] = lower ? x1 : x2;
        }
    }

    // ---- 5b. in-lane CR back-substitution --------------------------------
    const T xprev = lane_read_or_zero(xv[M - 1], lane - 1, ll >= 1);
#pragma unroll
    for (int k = KIN; k >= 1; --k) {
        const int SS = 1 << k, H = SS >> 1;
#pragma unroll
        for (int j = 0; j < M / SS; ++j) {
            const int i = SS * j + H - 1;
            const T xl = (i - H >= 0) ? xv[(i - H >= 0) ? i - H : 0] : xprev;
            xv[i] = backsub(e[i], xl, xv[i + H]);
        }
    }

    if constexpr (M % E == 0 && IO >= 2) {
        if (vec) {
            store_staged<T, M, IO == 3>(X, wave_base, total, lane, lds, xv);
            return;
        }
    } else if constexpr (M % E == 0) {
        if (vec) {
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
#define TRID_REG_CASE(RR)                                                                    \
    case RR:                                                                                 \
        if constexpr (RR <= M) {                                                             \
            hipLaunchKernelGGL((reg_crpcr_kernel<T, M, RR, IO>), grid, block, lds, g.stream, \
                               g.a, g.b, g.c, g.d, g.x, g.n, g.N, g.L, logL, g.rest, g.vec); \
            return hipGetLastError();                                                        \
        }                                                                                    \
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
    return hipErrorInvalidValue;
}

template <typename T, int M>
static hipError_t launch_reg_m(const RegArgs<T>& g, int R, dim3 grid, dim3 block)
{
    switch (g.io) {
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
