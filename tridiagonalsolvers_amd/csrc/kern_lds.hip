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
// * once 64 or fewer equations remain, wave 0 takes them into registers and
//   finishes the CR levels / the PCR on `rest` / the 2x2 solves with
//   cross-lane moves and no barrier, then writes the solved unknowns back;
// * a `rest` larger than 64 runs its PCR levels in place in LDS (registers
//   hold the new values between the two barriers of a level);
// * x overwrites d in LDS (d_i is dead once x_i is known), so four arrays of
//   n elements fit: FP32 up to n = 8192, FP64 up to n = 4096 in 160 KiB;
// * neighbours outside the system are identity rows, so any n is padded to a
//   power of two exactly.
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

// KPT: most PCR equations one thread updates per level when rest > 64.
template <typename T, int KPT>
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
    for (int i = tid; i < n_pad; i += nt) {
        Eq<T> e = identity_eq<T>();
        if (i < n) e = Eq<T>{A[base + i], B[base + i], C[base + i], D[base + i]};
        if (i == 0) e.a = T(0);          // a[first] ignored (tridiagonal_gold.h:28)
        if (i == n - 1) e.c = T(0);      // c[last] ignored
        s.put(i, e);
    }
    __syncthreads();

    // ---- CR forward levels in LDS while more than max(rest, 64) remain ------
    int st = 1;           // stride of the remaining equations (index st*k + st - 1)
    int count = n_pad;
    const int lds_stop = rest > kWave ? rest : kWave;
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

    if (count <= kWave) {
        // ---- wave tail: `count` equations, one per lane of wave 0 ----------
        if (tid < kWave) {
            const int L = count;
            const int ll = lane & (L - 1);
            const int idx = st * ll + st - 1;
            Eq<T> t = s.get(idx);
            const int r = rest < L ? rest : L;
            int h2 = 1;
            for (int cnt = L; cnt > r; cnt >>= 1, h2 <<= 1) {
                const bool active = ((ll + 1) & (2 * h2 - 1)) == 0;
                const Eq<T> lft = lane_read_eq(t, lane - h2, ll - h2 >= 0);
                const Eq<T> rgt = lane_read_eq(t, lane + h2, ll + h2 < L);
                const Eq<T> rr = reduce_eq(t, lft, rgt);
                if (active) t = rr;
            }
            const int LS = h2;
            const bool pact = ((ll + 1) & (LS - 1)) == 0;
            for (int dist = LS; dist < (L >> 1); dist <<= 1) {
                const Eq<T> lft = lane_read_eq(t, lane - dist, ll - dist >= 0);
                const Eq<T> rgt = lane_read_eq(t, lane + dist, ll + dist < L);
                const Eq<T> rr = reduce_eq(t, lft, rgt);
                if (pact) t = rr;
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
            if (lane < L) s.d[s.at(idx)] = xt;
        }
    } else {
        // ---- PCR on rest (> 64) equations in place in LDS ------------------
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
    while (st > 1) {
        const int h = st >> 1;
        for (int k = tid; k < n_pad / st; k += nt) {
            const int i = st * k + h - 1;
            const T xl = (i - h >= 0) ? s.d[s.at(i - h)] : T(0);
            const T xr = s.d[s.at(i + h)];
            const int p = s.at(i);
            s.d[p] = tdiv(s.d[p] - s.a[p] * xl - s.c[p] * xr, s.b[p]);
        }
        st = h;
        __syncthreads();
    }

    for (int i = tid; i < n; i += nt) X[base + i] = s.d[s.at(i)];
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
    int nt = n_pad / 2;
    if (nt > 256) nt = 256;
    if (nt < kWave) nt = kWave;
    const int kpt = g.rest > kWave ? (g.rest + nt - 1) / nt : 1;
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
    if (kpt <= 1) return go(lds_crpcr_kernel<T, 1>);
    if (kpt <= 4) return go(lds_crpcr_kernel<T, 4>);
    if (kpt <= 16) return go(lds_crpcr_kernel<T, 16>);
    return hipErrorInvalidValue;
}

template hipError_t launch_lds_crpcr<float>(const LdsArgs<float>&);
template hipError_t launch_lds_crpcr<double>(const LdsArgs<double>&);

}  // namespace trid
