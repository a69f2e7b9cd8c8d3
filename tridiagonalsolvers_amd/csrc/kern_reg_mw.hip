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
#include "reg_io.hpp"
#include "trid_device.hpp"
#include "trid_launch.hpp"

namespace trid {

template <typename T, int M, int IO>
__global__ __launch_bounds__(1024) void reg_mw_kernel(const T* __restrict__ A,
                                                      const T* __restrict__ B,
                                                      const T* __restrict__ C,
                                                      const T* __restrict__ D,
                                                      T* __restrict__ X, int64_t n, int rest,
                                                      int vec, int64_t N)
{
    constexpr int LOGM = Log2<M>::value;
    constexpr int E = 16 / (int)sizeof(T);
    constexpr int NPRE = LOGM + 6;  // published prefix equations per wave
    const int W = blockDim.x / kWave;
    const int w = threadIdx.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t sys = blockIdx.x;
    const int64_t base = sys * n;
    const int64_t j0 = (int64_t)(w * kWave + lane) * M;

    // LDS: the staging rows (load and store) share space with the exchange
    // area (prefix equations, reduced system, solved tops) used in between.
    extern __shared__ __align__(16) char trid_smem[];
    char* lds_stage = trid_smem + w * Stage<T, M>::kWaveBytes;
    Eq<T>* pre = reinterpret_cast<Eq<T>*>(trid_smem);   // [W][NPRE]
    Eq<T>* red = pre + W * NPRE;                          // [64*W]
    T* xs = reinterpret_cast<T*>(red + kWave * W);        // [64*W]

    // ---- load -------------------------------------------------------------
    Eq<T> e[M];
    bool loaded = false;
    if constexpr (M % E == 0 && IO >= 2) {
        if (vec) {
            load_staged<T, M, IO == 3>(A, B, C, D, base + (int64_t)w * kWave * M, n * N, lane,
                                       lds_stage, e);
            loaded = true;
        }
    } else if constexpr (M % E == 0) {
        if (vec) {
            load_direct<T, M, IO == 0>(A, B, C, D, base + j0, e);
            loaded = true;
        }
    }
    if (!loaded) {
#pragma unroll
        for (int s = 0; s < M; ++s) {
            const int64_t j = j0 + s;
            if (j < n) e[s] = Eq<T>{A[base + j], B[base + j], C[base + j], D[base + j]};
            else e[s] = identity_eq<T>();
        }
    }
    if (w == 0 && lane == 0) e[0].a = T(0);
#pragma unroll
    for (int s = 0; s < M; ++s)
        if (j0 + s == n - 1) e[s].c = T(0);
    const Eq<T> top0 = e[M - 1];  // level-0 top equation (lane 63's chain restarts here)
    __syncthreads();              // staging rows become the exchange area

    // ---- in-lane CR levels (lane 63's top slot is provisional) ------------
#pragma unroll
    for (int k = 1; k <= LOGM; ++k) {
        const int SS = 1 << k, H = SS >> 1;
        const Eq<T> nb = lane_read_eq(e[H - 1], lane + 1, lane + 1 < kWave);
#pragma unroll
        for (int j = 0; j < M / SS; ++j) {
            const int i = SS * j + SS - 1;
            const int ir = (i + H < M) ? i + H : i;
            const Eq<T> rgt = (i + H < M) ? e[ir] : nb;
            e[i] = reduce_eq(e[i], e[i - H], rgt);
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 1; k <= LOGM; ++k) pre[w * NPRE + k - 1] = e[(1 << (k - 1)) - 1];
    }

    // ---- cross-lane CR levels while more than max(rest, W) remain ---------
    int hs = 1;
    for (int count = kWave * W; count > rest && hs < kWave; count >>= 1, hs <<= 1) {
        const bool active = ((lane + 1) & (2 * hs - 1)) == 0;
        const Eq<T> lft = lane_read_eq(e[M - 1], lane - hs, lane - hs >= 0);
        const Eq<T> rgt = lane_read_eq(e[M - 1], lane + hs, lane + hs < kWave);
        const Eq<T> r = reduce_eq(e[M - 1], lft, rgt);
        if (active) e[M - 1] = r;
    }
    const int klane = __builtin_ctz((unsigned)hs);  // cross-lane levels done
    // prefix of cross-lane level j: lane 2^(j-1)-1 after level j-1 (final)
    if (((lane + 1) & lane) == 0 && lane + 1 <= (hs >> 1))
        pre[w * NPRE + LOGM + __builtin_ctz((unsigned)(lane + 1))] = e[M - 1];
    __syncthreads();

    // ---- redo lane 63's chain with the next wave's prefixes ---------------
    if (w + 1 < W) {
        const Eq<T>* nxt = pre + (w + 1) * NPRE;
        Eq<T> t = lane_bcast_eq(top0, kWave - 1);
#pragma unroll
        for (int k = 1; k <= LOGM; ++k) {
            const int H = 1 << (k - 1);
            t = reduce_eq(t, lane_bcast_eq(e[M - 1 - H], kWave - 1), nxt[k - 1]);
        }
        for (int j = 1; j <= klane; ++j) {
            const int h = 1 << (j - 1);
            t = reduce_eq(t, lane_bcast_eq(e[M - 1], kWave - 1 - h), nxt[LOGM + j - 1]);
        }
        if (lane == kWave - 1) e[M - 1] = t;
    }

    T xv[M];
#pragma unroll
    for (int s = 0; s < M; ++s) xv[s] = T(0);
    T xprevw;  // top unknown of the previous wave (0 for wave 0)

    if (rest <= W) {
        // ---- wave-top system: W equations, solved by wave 0 ---------------
        if (lane == kWave - 1) red[w] = e[M - 1];
        __syncthreads();
        if (w == 0) {
            const int L = W;
            const int ll = lane & (L - 1);
            Eq<T> t = red[ll];
            int h2 = 1;
            for (int count = L; count > rest; count >>= 1, h2 <<= 1) {
                const bool active = ((ll + 1) & (2 * h2 - 1)) == 0;
                const Eq<T> lft = lane_read_eq(t, lane - h2, ll - h2 >= 0);
                const Eq<T> rgt = lane_read_eq(t, lane + h2, ll + h2 < L);
                const Eq<T> r = reduce_eq(t, lft, rgt);
                if (active) t = r;
            }
            const int LS = h2;
            const bool pact = ((ll + 1) & (LS - 1)) == 0;
            for (int dist = LS; dist < (L >> 1); dist <<= 1) {
                const Eq<T> lft = lane_read_eq(t, lane - dist, ll - dist >= 0);
                const Eq<T> rgt = lane_read_eq(t, lane + dist, ll + dist < L);
                const Eq<T> r = reduce_eq(t, lft, rgt);
                if (pact) t = r;
            }
            const int half = L >> 1;
            const bool lower = ll < half;
            const Eq<T> partner = lane_read_eq(t, lower ? lane + half : lane - half, true);
            T x1, x2;
            solve2(lower ? t : partner, lower ? partner : t, x1, x2);
            T xt = lower ? x1 : x2;
            for (int h = LS >> 1; h >= 1; h >>= 1) {
                const T xl = lane_read_or_zero(xt, lane - h, ll - h >= 0);
                const T xr = lane_read_or_zero(xt, lane + h, ll + h < L);
                const bool solve = ((ll + 1) & (2 * h - 1)) == h;
                const T v = backsub(t, xl, xr);
                if (solve) xt = v;
            }
            if (lane < W) xs[lane] = xt;
        }
        __syncthreads();
        if (lane == kWave - 1) xv[M - 1] = xs[w];
        xprevw = w > 0 ? xs[w - 1] : T(0);
    } else {
        // ---- PCR on `rest` equations spread over the workgroup ------------
        const int LS = hs;  // lanes between consecutive PCR equations
        const bool pact = ((lane + 1) & (LS - 1)) == 0;
        const int q = (w * kWave + lane + 1) / LS - 1;
        if (pact) red[q] = e[M - 1];
        __syncthreads();
        for (int delta = 1; delta < (rest >> 1); delta <<= 1) {
            Eq<T> nw = e[M - 1];
            if (pact) {
                const Eq<T> lft = q - delta >= 0 ? red[q - delta] : identity_eq<T>();
                const Eq<T> rgt = q + delta < rest ? red[q + delta] : identity_eq<T>();
                nw = reduce_eq(e[M - 1], lft, rgt);
            }
            __syncthreads();
            if (pact) {
                red[q] = nw;
                e[M - 1] = nw;
            }
            __syncthreads();
        }
        if (pact) {
            const int half = rest >> 1;
            const bool lower = q < half;
            const Eq<T> partner = red[lower ? q + half : q - half];
            T x1, x2;
            solve2(lower ? e[M - 1] : partner, lower ? partner : e[M - 1], x1, x2);
            xv[M - 1] = lower ? x1 : x2;
            xs[q] = xv[M - 1];
        }
        __syncthreads();
        xprevw = w > 0 ? xs[(w * kWave) / LS - 1] : T(0);
    }

    // ---- cross-lane CR back-substitution ----------------------------------
    for (int h = hs >> 1; h >= 1; h >>= 1) {
        const T xl0 = lane_read(xv[M - 1], (lane - h) & (kWave - 1));
        const T xl = lane - h >= 0 ? xl0 : xprevw;
        const T xr = lane_read(xv[M - 1], (lane + h) & (kWave - 1));
        const bool solve = ((lane + 1) & (2 * h - 1)) == h;
        const T v = backsub(e[M - 1], xl, xr);
        if (solve) xv[M - 1] = v;
    }

    // ---- in-lane back-substitution ------------------------------------------
    const T xp0 = lane_read(xv[M - 1], (lane - 1) & (kWave - 1));
    const T xprev = lane >= 1 ? xp0 : xprevw;
#pragma unroll
    for (int k = LOGM; k >= 1; --k) {
        const int SS = 1 << k, H = SS >> 1;
#pragma unroll
        for (int j = 0; j < M / SS; ++j) {
            const int i = SS * j + H - 1;
            const T xl = (i - H >= 0) ? xv[(i - H >= 0) ? i - H : 0] : xprev;
            xv[i] = backsub(e[i], xl, xv[i + H]);
        }
    }

    // ---- store ----------------------------------------------------------------
    __syncthreads();  // exchange area becomes staging rows again
    if constexpr (M % E == 0 && IO >= 2) {
        if (vec) {
            store_staged<T, M, IO == 3>(X, base + (int64_t)w * kWave * M, n * N, lane, lds_stage,
                                        xv);
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

template <typename T, int M>
size_t mw_lds_bytes(int W, int io, int vec)
{
    constexpr int NPRE = Log2<M>::value + 6;
    const size_t xchg = (size_t)W * NPRE * sizeof(Eq<T>) + (size_t)kWave * W * sizeof(Eq<T>) +
                        (size_t)kWave * W * sizeof(T);
    const size_t stage = (io >= 2 && vec) ? (size_t)W * Stage<T, M>::kWaveBytes : 0;
    return xchg > stage ? xchg : stage;
}

template <typename T, int M, int IO>
static hipError_t launch_mw_io(const RegArgs<T>& g)
{
    const int W = g.L / kWave;
    const size_t lds = mw_lds_bytes<T, M>(W, IO, g.vec);
    if (lds > 160 * 1024) return hipErrorInvalidConfiguration;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)reg_mw_kernel<T, M, IO>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((reg_mw_kernel<T, M, IO>), dim3((unsigned)g.N), dim3(kWave * W), lds,
                       g.stream, g.a, g.b, g.c, g.d, g.x, g.n, g.rest, g.vec, g.N);
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
        case 16:
            if constexpr (sizeof(T) == 4) return launch_mw_m<T, 16>(g);
            return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_reg_mw<float>(const RegArgs<float>&);
template hipError_t launch_reg_mw<double>(const RegArgs<double>&);

}  // namespace trid
