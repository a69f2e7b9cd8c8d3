// kern_lds.hip -- LDS-resident CR / CR-PCR with a padded LDS map.
//
// Methods: TRID_CRNBC, TRID_CRPCRNBC (crNBC.cu), TRID_CRNBC32
// (crNBC32banks.cu), and the fallback for CR/CR-PCR/PCR requests the
// register kernels cannot take. As in the reference, one workgroup keeps a
// whole system in shared memory and runs the CR levels with a barrier each;
// unlike the reference:
//
// * the pad map is i -> i + (i >> log2(P)), one pad element per P elements.
//   At a CR level of stride s, the threads of one LDS access touch s*t + c;
//   with the map, t and t' collide on a bank only if s*(t - t') plus the
//   difference of their pad counts is a multiple of the bank count. A group
//   of B threads serviced together over B banks is conflict-free for every
//   s <= B when P = B (levels with larger s have at most n/s active
//   threads). Measured on gfx950 (SQ_LDS_BANK_CONFLICT, FP32 and FP64, pads
//   8..128 swept), P = 32 elements gives the fewest conflict cycles in both
//   precisions: 95% fewer than no padding in FP32, 81% fewer than P = 64 --
//   b32 accesses behave as 32 lanes over 32 banks at a time. The reference's
//   P = 16 (crNBC.cu:8-17, written for 16 banks) leaves 2-way conflicts.
//   P = 0 (no padding) is the crNBC32 control: its map (crNBC32banks.cu:8-17)
//   is the identity.
// * all loads of a thread are issued before the first LDS write (16-byte
//   vectors for dense, aligned batches), with at least one vector per thread
//   and array;
// * once 256 equations remain (64 below n = 256), wave 0 takes them into
//   registers, 4 consecutive ones per lane, and finishes the CR levels / the
//   PCR on `rest` / the 2x2 solves / the back-substitution of those levels
//   with the register core (reg_core.hpp: in-lane levels, DPP and
//   ds_bpermute moves, no barrier), then writes the solved unknowns back:
//   against a one-equation-per-lane tail at 64, two LDS levels and their
//   barriers fewer each way;
// * the LDS back-substitution takes two levels per barrier (the thread that
//   solves row i then solves rows i -+ h/2, whose neighbours it holds);
// * a `rest` larger than 256 runs its PCR levels in place in LDS (registers
//   hold the new values between the two barriers of a level);
// * x overwrites d in LDS (d_i is dead once x_i is known), so four arrays of
//   n elements fit: FP32 up to n = 8192, FP64 up to n = 4096 in 160 KiB;
// * neighbours outside the system are identity rows, so any n is padded to a
//   power of two exactly.
#include "reg_core.hpp"
#include "reg_io.hpp"
#include "trid_device.hpp"
#include "trid_launch.hpp"

namespace trid {

template <typename T>
struct LdsArrays {
    T* a;
    T* b;
    T* c;
    T* d;  // x overwrites d
    int sh;
    __device__ __forceinline__ int at(int i) const { return i + (i >> sh); }
    __device__ __forceinline__ Eq<T> get(int i) const
    {
        const int p = at(i);
        return Eq<T>{a[p], b[p], c[p], d[p]};
    }
    __device__ __forceinline__ void put(int i, const Eq<T>& e) const
    {
        const int p = at(i);
        a[p] = e.a;
        b[p] = e.b;
        c[p] = e.c;
        d[p] = e.d;
    }
};

// Equations handed to the register tail of wave 0 (kTailM per lane).
constexpr int kTailM = 4;
constexpr int kTail = kWave * kTailM;

// KPT: most PCR equations one thread updates per level when rest > 256.
// V > 0: dense, 16-byte aligned batch; each thread moves V 16-byte vectors
// per array (vector j = tid + nt*v holds elements [E*j, E*j + E)), and all of
// its loads are issued before the first LDS write. V = 0: element loop.
template <typename T, int KPT, int V>
__global__ __launch_bounds__(256) void lds_crpcr_kernel(const T* __restrict__ A,
                                                        const T* __restrict__ B,
                                                        const T* __restrict__ C,
                                                        const T* __restrict__ D,
                                                        T* __restrict__ X, int64_t n, int n_pad,
                                                        int rest, int sh)
{
    extern __shared__ __align__(16) char trid_smem[];
    const int np = n_pad + (n_pad >> sh);
    LdsArrays<T> s{reinterpret_cast<T*>(trid_smem), nullptr, nullptr, nullptr, sh};
    s.b = s.a + np;
    s.c = s.b + np;
    s.d = s.c + np;
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const int lane = tid & (kWave - 1);
    const int64_t base = (int64_t)blockIdx.x * n;
    lds_poison_wg(trid_smem, (unsigned)(4 * np * sizeof(T)));  // pad slots included

    // ---- load (coalesced), identity rows past n ---------------------------
    constexpr int E = 16 / (int)sizeof(T);
    if constexpr (V > 0) {
        u32x4 r[4][V];
        const T* src[4] = {A, B, C, D};
#pragma unroll
        for (int v = 0; v < V; ++v)
#pragma unroll
            for (int arr = 0; arr < 4; ++arr)
                r[arr][v] = gload16<true>(src[arr] + base + (int64_t)E * (tid + nt * v));
        T* dst[4] = {s.a, s.b, s.c, s.d};
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int i0 = E * (tid + nt * v);
#pragma unroll
            for (int arr = 0; arr < 4; ++arr) {
                T t[E];
                __builtin_memcpy(t, &r[arr][v], 16);
                if (arr == 0 && i0 == 0) t[0] = T(0);           // a[first] ignored
                if (arr == 2 && i0 + E == n) t[E - 1] = T(0);   // c[last] ignored
#pragma unroll
                for (int q = 0; q < E; ++q) dst[arr][s.at(i0 + q)] = t[q];
            }
        }
    } else {
        for (int i = tid; i < n_pad; i += nt) {
            Eq<T> e = identity_eq<T>();
            if (i < n) e = Eq<T>{A[base + i], B[base + i], C[base + i], D[base + i]};
            if (i == 0) e.a = T(0);          // a[first] ignored (tridiagonal_gold.h:28)
            if (i == n - 1) e.c = T(0);      // c[last] ignored
            s.put(i, e);
        }
    }
    __syncthreads();

    // ---- CR forward levels in LDS while more than max(rest, 256) remain -----
    int st = 1;           // stride of the remaining equations (index st*k + st - 1)
    int count = n_pad;
    // (systems shorter than kTail stop at 64 equations, one per lane, so that
    // every form keeps CR-PCR's split at `rest`)
    const int lds_stop = rest > kTail ? rest : (n_pad >= kTail ? kTail : kWave);
    while (count > lds_stop) {
        const int st2 = st << 1;
        for (int k = tid; k < (count >> 1); k += nt) {
            const int i = st2 * k + st2 - 1;
            const Eq<T> lft = s.get(i - st);
            const Eq<T> rgt = (i + st < n_pad) ? s.get(i + st) : identity_eq<T>();
            s.put(i, reduce_eq(s.get(i), lft, rgt));
        }
        st = st2;
        count >>= 1;
        __syncthreads();
    }

    if (count <= kTail) {
        // ---- wave tail: the last `count` equations in the registers of wave
        // 0 (kTailM consecutive ones per lane when count = kTail, else one
        // per lane of a segment of count lanes), solved by the register
        // CR / CR-PCR core: in-lane levels, DPP and ds_bpermute moves, no
        // barrier -------------------------------------------------------
        if (tid < kWave && count == kTail) {
            Eq<T> e[kTailM];
#pragma unroll
            for (int m = 0; m < kTailM; ++m) e[m] = s.get(st * (lane * kTailM + m) + st - 1);
            T xv[kTailM];
            if (rest <= kWave) wave_core<T, kTailM, 1, true>(e, xv, lane, lane, kWave, rest);
            else if (rest == 2 * kWave) wave_core<T, kTailM, 2, true>(e, xv, lane, lane, kWave, rest);
            else wave_core<T, kTailM, kTailM, true>(e, xv, lane, lane, kWave, rest);
#pragma unroll
            for (int m = 0; m < kTailM; ++m) s.d[s.at(st * (lane * kTailM + m) + st - 1)] = xv[m];
        } else if (tid < kWave) {
            const int L = count;
            const int ll = lane & (L - 1);
            const int idx = st * ll + st - 1;
            Eq<T> e[1] = {s.get(idx)};
            T xv[1];
            wave_core<T, 1, 1>(e, xv, lane, ll, L, rest < L ? rest : L);
            if (lane < L) s.d[s.at(idx)] = xv[0];
        }
    } else {
        // ---- PCR on rest (> 256) equations in place in LDS -----------------
        for (int delta = 1; delta < (rest >> 1); delta <<= 1) {
            Eq<T> nw[KPT];
#pragma unroll
            for (int j = 0; j < KPT; ++j) {
                const int k = tid + j * nt;
                if (k < rest) {
                    const int i = st * k + st - 1;
                    const Eq<T> lft =
                        k - delta >= 0 ? s.get(i - st * delta) : identity_eq<T>();
                    const Eq<T> rgt =
                        k + delta < rest ? s.get(i + st * delta) : identity_eq<T>();
                    nw[j] = reduce_eq(s.get(i), lft, rgt);
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < KPT; ++j) {
                const int k = tid + j * nt;
                if (k < rest) s.put(st * k + st - 1, nw[j]);
            }
            __syncthreads();
        }
        const int half = rest >> 1;
        T xv[KPT];
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
            const int k = tid + j * nt;
            if (k < rest) {
                const bool lower = k < half;
                const Eq<T> own = s.get(st * k + st - 1);
                const Eq<T> partner = s.get(st * (lower ? k + half : k - half) + st - 1);
                T x1, x2;
                solve2(lower ? own : partner, lower ? partner : own, x1, x2);
                xv[j] = lower ? x1 : x2;
            }
        }
        __syncthreads();  // all pairs read before any d slot becomes x
#pragma unroll
        for (int j = 0; j < KPT; ++j) {
            const int k = tid + j * nt;
            if (k < rest) s.d[s.at(st * k + st - 1)] = xv[j];
        }
    }
    __syncthreads();

    // ---- CR back-substitution in LDS ----------------------------------------
    // Two levels per barrier: the thread that solves row i at level h (from
    // x at i - h and i + h) then solves its two rows i -+ h/2 of the next
    // level, whose neighbours are x_i and those same two values.
    auto solve_row = [&](int i, T xl, T xr) {
        const int p = s.at(i);
        return tdiv(s.d[p] - s.a[p] * xl - s.c[p] * xr, s.b[p]);
    };
    while (st > 1) {
        const int h = st >> 1, q = st >> 2;
        for (int k = tid; k < n_pad / st; k += nt) {
            const int i = st * k + h - 1;
            const T xl = (i - h >= 0) ? s.d[s.at(i - h)] : T(0);
            const T xr = s.d[s.at(i + h)];
            const T xi = solve_row(i, xl, xr);
            if (q > 0) {
                const T xa = solve_row(i - q, xl, xi);
                const T xb = solve_row(i + q, xi, xr);
                s.d[s.at(i - q)] = xa;
                s.d[s.at(i + q)] = xb;
            }
            s.d[s.at(i)] = xi;
        }
        st = q > 0 ? q : h;
        __syncthreads();
    }

    if constexpr (V > 0) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int i0 = E * (tid + nt * v);
            T t[E];
#pragma unroll
            for (int q = 0; q < E; ++q) t[q] = s.d[s.at(i0 + q)];
            u32x4 w;
            __builtin_memcpy(&w, t, 16);
            gstore16<true>(X + base + i0, w);
        }
    } else {
        for (int i = tid; i < n; i += nt) X[base + i] = s.d[s.at(i)];
    }
}

template <typename T>
hipError_t launch_lds_crpcr(const LdsArgs<T>& g)
{
    if (g.N <= 0 || g.N > 0x7fffffffLL) return hipErrorInvalidConfiguration;
    const int n_pad = g.n_pad;
    const int sh = g.pad > 0 ? __builtin_ctz((unsigned)g.pad) : 31;
    const int np = n_pad + (n_pad >> sh);
    const size_t lds = (size_t)4 * np * sizeof(T);
    if (lds > 160 * 1024) return hipErrorInvalidConfiguration;
    // threads per system: at least one 16-byte vector per array each, so that
    // dense batches take the vector loads (131,072 x 512 FP32 crNBC: 0.24 ms
    // at 128 threads, 0.39 ms at 256 threads and element loads); at most 256
    constexpr int E = 16 / (int)sizeof(T);
    int nt = n_pad / E;
    if (nt > 256) nt = 256;
    if (nt < kWave) nt = kWave;
    const int kpt = g.rest > kTail ? (g.rest + nt - 1) / nt : 1;
    auto al16 = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; };
    int V = 0;
    if (g.n == n_pad && n_pad >= E * nt && al16(g.a) && al16(g.b) && al16(g.c) && al16(g.d) &&
        al16(g.x))
        V = n_pad / (E * nt);
    if (V > 8) V = 0;
    auto go = [&](auto kern) -> hipError_t {
        if (lds > 64 * 1024) {
            hipError_t e = hipFuncSetAttribute((const void*)kern,
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)lds);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(kern, dim3((unsigned)g.N), dim3(nt), lds, g.stream, g.a, g.b, g.c,
                           g.d, g.x, g.n, n_pad, g.rest, sh);
        return hipGetLastError();
    };
    auto by_v = [&](auto kpt_tag) -> hipError_t {
        constexpr int K = decltype(kpt_tag)::value;
        switch (V) {
            case 1: return go(lds_crpcr_kernel<T, K, 1>);
            case 2: return go(lds_crpcr_kernel<T, K, 2>);
            case 4: return go(lds_crpcr_kernel<T, K, 4>);
            case 8: return go(lds_crpcr_kernel<T, K, 8>);
            default: return go(lds_crpcr_kernel<T, K, 0>);
        }
    };
    if (kpt <= 1) return by_v(std::integral_constant<int, 1>{});
    if (kpt <= 4) return by_v(std::integral_constant<int, 4>{});
    if (kpt <= 16) return by_v(std::integral_constant<int, 16>{});
    return hipErrorInvalidValue;
}

template hipError_t launch_lds_crpcr<float>(const LdsArgs<float>&);
template hipError_t launch_lds_crpcr<double>(const LdsArgs<double>&);

}  // namespace trid
