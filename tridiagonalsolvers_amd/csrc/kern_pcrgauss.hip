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
//   launch on tiles with a halo of G-1 equations on each side, held in LDS;
//   pass 2 runs Thomas with one lane per subsystem. Lanes r, r+1, ... read
//   consecutive addresses at every step, so pass 2 is coalesced; its
//   forward-sweep factors c', d' overwrite the pass-1 workspace in place.
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

// Pass 1: log2(G) PCR levels on tiles of `tile` equations with a halo of
// G-1 on each side. Output (a', b', c', d') goes to the workspace arrays.
template <typename T, int KPT>
__global__ __launch_bounds__(512) void pcr_global_levels_kernel(
    const T* __restrict__ A, const T* __restrict__ B, const T* __restrict__ C,
    const T* __restrict__ D, T* __restrict__ A2, T* __restrict__ B2, T* __restrict__ C2,
    T* __restrict__ D2, int64_t n, int G, int tile)
{
    extern __shared__ __align__(16) char trid_smem[];
    const int H = G - 1;
    const int len = tile + 2 * H;
    T* sa = reinterpret_cast<T*>(trid_smem);
    T* sb = sa + len;
    T* sc = sb + len;
    T* sd = sc + len;
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const int64_t sys = blockIdx.y;
    const int64_t base = sys * n;
    const int64_t t0 = (int64_t)blockIdx.x * tile;

    for (int p = tid; p < len; p += nt) {
        const int64_t g = t0 - H + p;
        Eq<T> e = identity_eq<T>();
        if (g >= 0 && g < n) e = Eq<T>{A[base + g], B[base + g], C[base + g], D[base + g]};
        if (g == 0) e.a = T(0);
        if (g == n - 1) e.c = T(0);
        sa[p] = e.a;
        sb[p] = e.b;
        sc[p] = e.c;
        sd[p] = e.d;
    }
    __syncthreads();
    for (int delta = 1; delta < G; delta <<= 1) {
        Eq<T> nw[KPT];
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
            const int p = tid + j * nt;
            if (p < len) {
                const Eq<T> own{sa[p], sb[p], sc[p], sd[p]};
                // outside the loaded window the value is never used by the
                // tile (the window shrinks by delta per level, G-1 in total)
                const Eq<T> lft = p - delta >= 0
                                      ? Eq<T>{sa[p - delta], sb[p - delta], sc[p - delta], sd[p - delta]}
                                      : identity_eq<T>();
                const Eq<T> rgt = p + delta < len
                                      ? Eq<T>{sa[p + delta], sb[p + delta], sc[p + delta], sd[p + delta]}
                                      : identity_eq<T>();
                nw[j] = reduce_eq(own, lft, rgt);
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
            const int p = tid + j * nt;
            if (p < len) {
                // rows outside the system stay identity rows
                const int64_t g = t0 - H + p;
                if (g >= 0 && g < n) {
                    sa[p] = nw[j].a;
                    sb[p] = nw[j].b;
                    sc[p] = nw[j].c;
                    sd[p] = nw[j].d;
                }
            }
        }
        __syncthreads();
    }
    for (int p = tid; p < tile; p += nt) {
        const int64_t g = t0 + p;
        if (g < n) {
            A2[base + g] = sa[H + p];
            B2[base + g] = sb[H + p];
            C2[base + g] = sc[H + p];
            D2[base + g] = sd[H + p];
        }
    }
}

// Pass 2: Thomas on subsystem r of system s, one lane each. The forward
// factors overwrite C2/D2; x goes to X.
template <typename T>
__global__ __launch_bounds__(256) void thomas_strided_kernel(const T* __restrict__ A2,
                                                             const T* __restrict__ B2,
                                                             T* __restrict__ C2,
                                                             T* __restrict__ D2,
                                                             T* __restrict__ X, int64_t n,
                                                             int64_t N, int G)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= N * G) return;
    const int64_t s = gid / G;
    const int r = (int)(gid % G);
    const int64_t base = s * n;
    if (r >= n) return;
    const int64_t m = (n - r + G - 1) / G;  // elements r, r+G, ... below n
    constexpr int U = 8;                      // loads issued ahead
    int64_t i = base + r;
    T cp = tdiv(C2[i], B2[i]);
    T dp = tdiv(D2[i], B2[i]);
    C2[i] = cp;
    D2[i] = dp;
    int64_t j = 1;
    for (; j + U <= m; j += U) {
        T ta[U], tb[U], tc[U], td[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t q = base + r + (j + u) * G;
            ta[u] = A2[q];
            tb[u] = B2[q];
            tc[u] = C2[q];
            td[u] = D2[q];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const T mm = tb[u] - ta[u] * cp;
            cp = tdiv(tc[u], mm);
            dp = tdiv(td[u] - dp * ta[u], mm);
            tc[u] = cp;
            td[u] = dp;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t q = base + r + (j + u) * G;
            C2[q] = tc[u];
            D2[q] = td[u];
        }
    }
    for (; j < m; ++j) {
        const int64_t q = base + r + j * G;
        const T ai = A2[q];
        const T mm = B2[q] - ai * cp;
        cp = tdiv(C2[q], mm);
        dp = tdiv(D2[q] - dp * ai, mm);
        C2[q] = cp;
        D2[q] = dp;
    }
    T xv = dp;
    X[base + r + (m - 1) * G] = xv;
    j = m - 2;
    for (; j - U + 1 >= 0; j -= U) {
        T tc[U], td[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t q = base + r + (j - u) * G;
            tc[u] = C2[q];
            td[u] = D2[q];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            xv = td[u] - tc[u] * xv;
            td[u] = xv;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) X[base + r + (j - u) * G] = td[u];
    }
    for (; j >= 0; --j) {
        const int64_t q = base + r + j * G;
        xv = D2[q] - C2[q] * xv;
        X[q] = xv;
    }
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
    int tile = 4096;
    if (tile < 4 * G) tile = 4 * G;
    const int len = tile + 2 * (G - 1);
    const int nt = 512;
    const int kpt = (len + nt - 1) / nt;
    const size_t lds = (size_t)4 * len * sizeof(T);
    if (lds > 160 * 1024) return hipErrorInvalidConfiguration;
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
        if (kpt <= 10) e = go(pcr_global_levels_kernel<T, 10>);
        else if (kpt <= 12 && sizeof(T) == 4) e = go(pcr_global_levels_kernel<T, 12>);
        else return hipErrorInvalidValue;
        if (e != hipSuccess) return e;
    }
    const int64_t lanes = g.N * (int64_t)G;
    const int64_t blocks = (lanes + 255) / 256;
    hipLaunchKernelGGL((thomas_strided_kernel<T>), dim3((unsigned)blocks), dim3(256), 0,
                       g.stream, A2, B2, C2, D2, g.x, g.n, g.N, G);
    return hipGetLastError();
}

template hipError_t launch_pcr_thomas<float>(const PcrThomasArgs<float>&);
template hipError_t launch_pcr_thomas<double>(const PcrThomasArgs<double>&);

}  // namespace trid
