// kern_pcrgauss.hip -- PCR to G interleaved subsystems, then Thomas on each.
//
// Method: TRID_PCRGAUSS (pcrGauss_kernel.cu, pcrGauss_app.cu). After
// log2(G) PCR levels (distances 1, 2, ..., G/2) equation i couples only to
// i +- G, so the system splits exactly into G subsystems {r, r+G, r+2G, ...}
// that Thomas solves independently. The PCR depth is always exactly log2(G);
// the reference's template tables pair some G with a different depth
// (pcrGauss_app.cu:24-45,71-94), which is only approximately right.
//
// Two forms:
// * in LDS (n up to 8192 FP32 / 4096 FP64): one workgroup per system, the
//   PCR levels in place in padded LDS (new values held in registers between
//   the two barriers of a level), then one thread per subsystem walks it with
//   Thomas. Threads r, r+1, ... touch consecutive LDS words at every step, so
//   the Thomas sweeps are conflict-free.
// * two passes through HBM for long systems (the reference's large path,
//   pcrGauss_app.cu:102-162, runs one global PCR level per launch and keeps
//   one 1024-thread block per system): pass 1 runs all log2(G) levels in one
//   launch on tiles with a halo of G-1 equations on each side, held in
//   registers and exchanged through LDS; pass 2 runs Thomas with one lane per
//   subsystem. Lanes r, r+1, ... read consecutive addresses at every step,
//   so pass 2 is coalesced; its forward-sweep factors c', d' overwrite the
//   pass-1 workspace in place.
#include "trid_device.hpp"
#include "trid_launch.hpp"

namespace trid {

// ---------------------------------------------------------------------------
// in-LDS form
// ---------------------------------------------------------------------------

template <typename T, int KPT>
__global__ __launch_bounds__(1024) void pcr_thomas_lds_kernel(const T* __restrict__ A,
                                                              const T* __restrict__ B,
                                                              const T* __restrict__ C,
                                                              const T* __restrict__ D,
                                                              T* __restrict__ X, int64_t n,
                                                              int n_pad, int G, int sh)
{
    extern __shared__ __align__(16) char trid_smem[];
    const int np = n_pad + (n_pad >> sh);
    T* sa = reinterpret_cast<T*>(trid_smem);
    T* sb = sa + np;
    T* sc = sb + np;
    T* sd = sc + np;
    auto at = [sh](int i) { return i + (i >> sh); };
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const int64_t base = (int64_t)blockIdx.x * n;

    for (int i = tid; i < n_pad; i += nt) {
        Eq<T> e = identity_eq<T>();
        if (i < n) e = Eq<T>{A[base + i], B[base + i], C[base + i], D[base + i]};
        if (i == 0) e.a = T(0);
        if (i == n - 1) e.c = T(0);
        const int p = at(i);
        sa[p] = e.a;
        sb[p] = e.b;
        sc[p] = e.c;
        sd[p] = e.d;
    }
    __syncthreads();

    // log2(G) PCR levels (pcrGauss_kernel.cu:82-110 with depth = log2 G)
    for (int delta = 1; delta < G; delta <<= 1) {
        Eq<T> nw[KPT];
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
            const int i = tid + j * nt;
            if (i < n_pad) {
                const int p = at(i);
                const Eq<T> own{sa[p], sb[p], sc[p], sd[p]};
                Eq<T> lft = identity_eq<T>(), rgt = identity_eq<T>();
                if (i - delta >= 0) {
                    const int q = at(i - delta);
                    lft = Eq<T>{sa[q], sb[q], sc[q], sd[q]};
                }
                if (i + delta < n_pad) {
                    const int q = at(i + delta);
                    rgt = Eq<T>{sa[q], sb[q], sc[q], sd[q]};
                }
                nw[j] = reduce_eq(own, lft, rgt);
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
            const int i = tid + j * nt;
            if (i < n_pad) {
                const int p = at(i);
                sa[p] = nw[j].a;
                sb[p] = nw[j].b;
                sc[p] = nw[j].c;
                sd[p] = nw[j].d;
            }
        }
        __syncthreads();
    }

    // Thomas on subsystem r (tridiagonal_gold.h:26-44 order of operations);
    // c' and d' overwrite c and d, then x overwrites d.
    const int m = n_pad / G;
    for (int r = tid; r < G; r += nt) {
        int p = at(r);
        T cp = tdiv(sc[p], sb[p]);
        T dp = tdiv(sd[p], sb[p]);
        sc[p] = cp;
        sd[p] = dp;
        for (int j = 1; j < m; ++j) {
            p = at(r + j * G);
            const T ai = sa[p];
            const T mm = sb[p] - ai * cp;
            cp = tdiv(sc[p], mm);
            dp = tdiv(sd[p] - dp * ai, mm);
            sc[p] = cp;
            sd[p] = dp;
        }
        T xv = dp;
        for (int j = m - 2; j >= 0; --j) {
            p = at(r + j * G);
            xv = sd[p] - sc[p] * xv;
            sd[p] = xv;
        }
    }
    __syncthreads();
    for (int i = tid; i < n; i += nt) X[base + i] = sd[at(i)];
}

// ---------------------------------------------------------------------------
// two-pass form
// ---------------------------------------------------------------------------

// Tiles of the global PCR pass: at least TRID_PCRG_TILE_MIN equations and
// TRID_PCRG_TILE_PER_G * G (so the halo of 2(G-1) stays a small fraction),
// halved while the window does not fit the registers (12 equations per
// thread FP32, 10 FP64) or LDS; TRID_PCRG_THREADS threads per workgroup.
// Measured on 64 x 2^20 FP32 (depth 2..10): 2048 / 8 G / 512 threads is
// the best of the shapes tried (tiles of 2048..8192 equations, 256..1024
// threads).
#ifndef TRID_PCRG_TILE_MIN
#define TRID_PCRG_TILE_MIN 2048
#endif
#ifndef TRID_PCRG_TILE_PER_G
#define TRID_PCRG_TILE_PER_G 8
#endif
#ifndef TRID_PCRG_THREADS
#define TRID_PCRG_THREADS 512
#endif

// Pass 1: log2(G) PCR levels on tiles of `tile` equations with a halo of
// G-1 on each side. Each thread keeps its equations (p = tid + j*nt of the
// window) in registers; a level publishes them to LDS as one record per
// equation, and reads the two neighbours back (48 bytes of LDS traffic per
// equation and level for FP32). Output (a', b', c', d') goes to the
// workspace arrays.
template <typename T, int KPT>
__global__ __launch_bounds__(TRID_PCRG_THREADS) void pcr_global_levels_kernel(
    const T* __restrict__ A, const T* __restrict__ B, const T* __restrict__ C,
    const T* __restrict__ D, T* __restrict__ A2, T* __restrict__ B2, T* __restrict__ C2,
    T* __restrict__ D2, int64_t n, int G, int tile)
{
    extern __shared__ __align__(16) char trid_smem[];
    const int H = G - 1;
    const int len = tile + 2 * H;
    Eq<T>* se = reinterpret_cast<Eq<T>*>(trid_smem);
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const int64_t sys = blockIdx.y;
    const int64_t base = sys * n;
    const int64_t t0 = (int64_t)blockIdx.x * tile;

    Eq<T> own[KPT];
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
        const int p = tid + j * nt;
        const int64_t g = t0 - H + p;
        own[j] = identity_eq<T>();
        if (p < len && g >= 0 && g < n) {
            own[j] = Eq<T>{A[base + g], B[base + g], C[base + g], D[base + g]};
            if (g == 0) own[j].a = T(0);
            if (g == n - 1) own[j].c = T(0);
        }
    }
    for (int delta = 1; delta < G; delta <<= 1) {
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
            const int p = tid + j * nt;
            if (p < len) se[p] = own[j];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
            const int p = tid + j * nt;
            const int64_t g = t0 - H + p;
            // rows outside the system stay identity rows; outside the loaded
            // window the value is never used by the tile (the window shrinks
            // by delta per level, G-1 in total)
            if (p < len && g >= 0 && g < n) {
                const Eq<T> lft = p - delta >= 0 ? se[p - delta] : identity_eq<T>();
                const Eq<T> rgt = p + delta < len ? se[p + delta] : identity_eq<T>();
                own[j] = reduce_eq(own[j], lft, rgt);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
        const int p = tid + j * nt;
        const int64_t g = t0 - H + p;
        if (p >= H && p < H + tile && g < n) {
            A2[base + g] = own[j].a;
            B2[base + g] = own[j].b;
            C2[base + g] = own[j].c;
            D2[base + g] = own[j].d;
        }
    }
}

// Strided elimination (pass 2 of the long form, and the interleaved-batch
// solve): chain r of block s is the rows  base(s) + r + k * stride,
// k = 0 .. m-1, base(s) = s * block_len, r < chains, m = the rows below
// block_len. For pcrGauss the blocks are the systems and the chains their G
// interleaved subsystems (stride G); an interleaved batch is one block whose
// chains are the systems (row i of system j at i * stride + j).
//
// Each chain is eliminated from both ends at once (a twisted
// factorization): lane l < 32 of a wave eliminates rows 0 .. h-1 downwards
// (the Thomas forward sweep), lane l + 32 eliminates rows m-1 .. h upwards,
// the two meet in a 2x2 system for x_{h-1}, x_h, and each lane
// back-substitutes its half. This halves the serial chain of Thomas without
// changing the elimination it performs on each half. Half-waves r, r+1, ...
// read consecutive addresses at every step. The chains are long and few, so
// the kernel runs one wave per workgroup (spreading them over every CU) and
// keeps the next U steps' loads in flight while it computes the current U
// (double buffering). Loads and stores share one 6-bit outstanding-operation
// counter, so waiting for a batch while the next batch's 4U loads and this
// batch's 2U stores are in flight needs 6U <= 63: U = 8.
//
// In elimination order i = 0, 1, ... a lane visits row k_i (k_i = i on the
// top half, m-1-i on the bottom half) with p the coupling to the row
// already eliminated (a on top, c below) and q the coupling to the next
// (c on top, a below):
//   den = b - p q'_{i-1},  q'_i = q / den,  d'_i = (d - p d'_{i-1}) / den
// (p ignored for i = 0), so x_{k_i} = d'_i - q'_i x_{k_{i+1}}. At the meeting
// point  x_{h-1} + c' x_h = d'_top  and  x_h + a' x_{h-1} = d'_bottom.
// The factors q', d' go to FC, FD (which may be C, D themselves, or X for
// FD: each location is read before the same lane overwrites it).
template <typename T, int U>
__global__ __launch_bounds__(64) void thomas_strided_kernel(const T* A, const T* B, const T* C,
                                                            const T* D, T* FC, T* FD, T* X,
                                                            int64_t block_len, int64_t blocks,
                                                            int64_t chains, int64_t stride)
{
    const int lane = threadIdx.x;
    const bool top = lane < kWave / 2;
    const int64_t chain = (int64_t)blockIdx.x * (kWave / 2) + (lane & (kWave / 2 - 1));
    const int64_t s = chain / chains;
    const int64_t r = chain % chains;
    const bool live = chain < blocks * chains && r < block_len;
    const int64_t m = live ? (block_len - r + stride - 1) / stride : 0;
    const int64_t h = (m + 1) / 2;  // rows of the top half
    const int64_t len = top ? h : m - h;
    // row k_i = first + i * step
    const int64_t first = top ? 0 : m - 1;
    const int64_t step = top ? stride : -stride;
    const int64_t o0 = s * block_len + r + first * stride;
    const T* pp = (top ? A : C) + o0;  // coupling to the eliminated side
    const T* pb = B + o0;
    const T* pq = (top ? C : A) + o0;  // coupling to the next row
    const T* pd = D + o0;
    T* wq = FC + o0;
    T* wd = FD + o0;
    T* px = X + o0;

    T qv = T(0), dv = T(0);
    T np[U], nb[U], nq[U], nd[U];
    auto fetch_elim = [&](int64_t i0) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = i0 + u;
            if (i < len) {
                np[u] = pp[i * step];
                nb[u] = pb[i * step];
                nq[u] = pq[i * step];
                nd[u] = pd[i * step];
            }
        }
    };
    fetch_elim(0);
    for (int64_t i0 = 0; i0 < len; i0 += U) {
        T tp[U], tb[U], tq[U], td[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            tp[u] = np[u];
            tb[u] = nb[u];
            tq[u] = nq[u];
            td[u] = nd[u];
        }
        fetch_elim(i0 + U);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = i0 + u;
            if (i < len) {
                const T pk = i > 0 ? tp[u] : T(0);
                const T den = tb[u] - pk * qv;
                qv = tdiv(tq[u], den);
                dv = tdiv(td[u] - dv * pk, den);
                wq[i * step] = qv;
                wd[i * step] = dv;
            }
        }
    }
    // meeting point: lane l and l + 32 hold the two normalized rows
    const int partner = top ? lane + kWave / 2 : lane - kWave / 2;
    const T oq = lane_read(qv, partner);
    const T od = lane_read(dv, partner);
    T xv;
    if (top) {
        // x_{h-1} + qv x_h = dv;  x_h + oq x_{h-1} = od (absent when m == 1)
        xv = m - h > 0 ? tdiv(dv - qv * od, T(1) - qv * oq) : dv;
    } else {
        const T xo = tdiv(od - oq * dv, T(1) - oq * qv);  // x_{h-1}
        xv = dv - qv * xo;                                // x_h
    }
    if (len > 0) px[(len - 1) * step] = xv;
    // back-substitution, i = len-2 .. 0
    auto fetch_back = [&](int64_t i0) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = i0 - u;
            if (i >= 0) {
                nq[u] = wq[i * step];
                nd[u] = wd[i * step];
            }
        }
    };
    fetch_back(len - 2);
    for (int64_t i0 = len - 2; i0 >= 0; i0 -= U) {
        T tq[U], td[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            tq[u] = nq[u];
            td[u] = nd[u];
        }
        fetch_back(i0 - U);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = i0 - u;
            if (i >= 0) {
                xv = td[u] - tq[u] * xv;
                px[i * step] = xv;
            }
        }
    }
}

template <typename T>
static hipError_t launch_strided(const T* A, const T* B, const T* C, const T* D, T* FC, T* FD,
                                 T* X, int64_t block_len, int64_t blocks, int64_t chains,
                                 int64_t stride, hipStream_t stream)
{
    const int64_t total = blocks * chains;  // two lanes each
    const int64_t grid = (total + kWave / 2 - 1) / (kWave / 2);
    if (total <= 0) return hipSuccess;
    if (grid > 0x7fffffffLL) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL((thomas_strided_kernel<T, 8>), dim3((unsigned)grid), dim3(kWave), 0, stream,
                       A, B, C, D, FC, FD, X, block_len, blocks, chains, stride);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_interleaved(const InterleavedArgs<T>& g)
{
    if (g.N <= 0) return hipSuccess;
    // one block of n * stride elements whose N chains are the systems; the
    // c' factors go to the workspace, the d' factors to x
    return launch_strided<T>(g.a, g.b, g.c, g.d, g.ws, g.x, g.x, g.n * g.stride, 1, g.N,
                             g.stride, g.stream);
}

template <typename T>
hipError_t launch_pcr_thomas(const PcrThomasArgs<T>& g)
{
    if (g.N <= 0) return hipSuccess;
    if (!g.two_pass) {
        const int n_pad = g.n_pad;
        const int sh = g.pad > 0 ? __builtin_ctz((unsigned)g.pad) : 31;
        const int np = n_pad + (n_pad >> sh);
        const size_t lds = (size_t)4 * np * sizeof(T);
        if (lds > 160 * 1024 || g.N > 0x7fffffffLL) return hipErrorInvalidConfiguration;
        int nt = n_pad < 1024 ? n_pad : 1024;
        if (nt < kWave) nt = kWave;
        const int kpt = (n_pad + nt - 1) / nt;
        auto go = [&](auto kern) -> hipError_t {
            if (lds > 64 * 1024) {
                hipError_t e = hipFuncSetAttribute(
                    (const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                if (e != hipSuccess) return e;
            }
            hipLaunchKernelGGL(kern, dim3((unsigned)g.N), dim3(nt), lds, g.stream, g.a, g.b,
                               g.c, g.d, g.x, g.n, n_pad, g.G, sh);
            return hipGetLastError();
        };
        if (kpt <= 1) return go(pcr_thomas_lds_kernel<T, 1>);
        if (kpt <= 2) return go(pcr_thomas_lds_kernel<T, 2>);
        if (kpt <= 4) return go(pcr_thomas_lds_kernel<T, 4>);
        if (kpt <= 8) return go(pcr_thomas_lds_kernel<T, 8>);
        return hipErrorInvalidValue;
    }
    // two passes through the workspace (4 arrays of n*N elements)
    const int64_t total = g.n * g.N;
    T* A2 = g.ws;
    T* B2 = A2 + total;
    T* C2 = B2 + total;
    T* D2 = C2 + total;
    const int G = g.G;
    const int nt = TRID_PCRG_THREADS;
    const int max_kpt = sizeof(T) == 4 ? 12 : 10;
    int tile = TRID_PCRG_TILE_MIN;
    if (tile < TRID_PCRG_TILE_PER_G * G) tile = TRID_PCRG_TILE_PER_G * G;
    while (tile > G && (tile + 2 * (G - 1) > max_kpt * nt ||
                        (size_t)(tile + 2 * (G - 1)) * sizeof(Eq<T>) > 160 * 1024))
        tile /= 2;
    const int len = tile + 2 * (G - 1);
    const int kpt = (len + nt - 1) / nt;
    const size_t lds = (size_t)len * sizeof(Eq<T>);
    if (lds > 160 * 1024 || kpt > max_kpt) return hipErrorInvalidConfiguration;
    const int64_t tiles = (g.n + tile - 1) / tile;
    if (tiles > 0x7fffffffLL || g.N > 65535) return hipErrorInvalidConfiguration;
    {
        auto go = [&](auto kern) -> hipError_t {
            if (lds > 64 * 1024) {
                hipError_t e = hipFuncSetAttribute(
                    (const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                if (e != hipSuccess) return e;
            }
            hipLaunchKernelGGL(kern, dim3((unsigned)tiles, (unsigned)g.N), dim3(nt), lds,
                               g.stream, g.a, g.b, g.c, g.d, A2, B2, C2, D2, g.n, G, tile);
            return hipGetLastError();
        };
        // G = 1 runs no PCR level: pass 1 only copies (and zeroes a[0], c[n-1])
        hipError_t e;
        if (kpt <= 4) e = go(pcr_global_levels_kernel<T, 4>);
        else if (kpt <= 6) e = go(pcr_global_levels_kernel<T, 6>);
        else if (kpt <= 8) e = go(pcr_global_levels_kernel<T, 8>);
        else if (kpt <= 10) e = go(pcr_global_levels_kernel<T, 10>);
        else if (kpt <= 12 && sizeof(T) == 4) e = go(pcr_global_levels_kernel<T, 12>);
        else return hipErrorInvalidValue;
        if (e != hipSuccess) return e;
    }
    return launch_strided<T>(A2, B2, C2, D2, C2, D2, g.x, g.n, g.N, G, G, g.stream);
}

template hipError_t launch_pcr_thomas<float>(const PcrThomasArgs<float>&);
template hipError_t launch_pcr_thomas<double>(const PcrThomasArgs<double>&);
template hipError_t launch_interleaved<float>(const InterleavedArgs<float>&);
template hipError_t launch_interleaved<double>(const InterleavedArgs<double>&);

}  // namespace trid
