// reg_core.hpp -- the register-resident CR / CR-PCR / PCR solve of one
// system held by a segment of L lanes (M consecutive unknowns per lane),
// shared by the batch kernel (kern_reg.hip), the refinement kernels
// (kern_refine.hip) and the ragged kernel (kern_ragged.hip).
//
//   1. in-lane CR levels      (stride 2..M; the top equation of a lane reads
//                              the right lane's "prefix" equation by bpermute)
//   2. cross-lane CR levels   (while more than `rest` equations remain)
//   3. PCR on the `rest` system, cross-lane, all reads before writes because
//      every level works on registers (R > 1: on rows scaled to a unit
//      diagonal, pcr_rows)
//   4. 2x2 Cramer solves at lane distance L/2
//   5. CR back-substitution across lanes, then inside lanes
//
// R: equations per lane handed to PCR (R = 1: rest <= L; R > 1: rest = R*L).
// FULL: the segment is the whole wave (L == 64), known at compile time.
// On return xv holds the lane's M unknowns; e holds the CR-reduced equations
// (for R > 1 its PCR slots keep their values from before PCR).
#pragma once

#include <type_traits>

#include "trid_device.hpp"

namespace trid {

// PCR on the rest = R * L system of a segment of L lanes: lane ll holds PCR
// equations q = ll*R + r at slots S*r + S - 1 (S = M / R), r = 0..R-1. The
// rows are scaled to a unit diagonal first (three values per row, one
// reciprocal per reduction). Every lane read stays inside the segment, so a
// read that lies outside the system returns a finite row of this system,
// which meets a coefficient that is exactly 0: no selects, and a non-finite
// system cannot reach the other systems of its wave. FULL (L == 64, known at
// compile time): one-lane steps are DPP wave shifts (the ends of the wave
// shift in zeros); otherwise every read is a ds_bpermute whose lane index
// wraps within the segment. On return the PCR slots of xv hold their
// unknowns.
template <typename T, int M, int R, bool FULL>
__device__ __forceinline__ void pcr_rows(Eq<T> (&e)[M], T (&xv)[M], int lane, int ll, int L)
{
    constexpr int S = M / R;
    if constexpr (FULL) {
        L = kWave;
        ll = lane;
    }
    NEq<T> q[R];
#pragma unroll
    for (int r = 0; r < R; ++r) q[r] = normalize_eq(e[S * r + S - 1]);
    const int seg = lane - ll;  // first lane of the segment
    auto src = [&](int off) { return seg + ((ll + off) & (L - 1)); };
    const int sm1 = src(-1), sp1 = src(1);
    auto left1 = [&](const NEq<T>& v) {
        if constexpr (FULL) return lane_step_neq<-1>(v);
        else return lane_read_neq(v, sm1);
    };
    auto right1 = [&](const NEq<T>& v) {
        if constexpr (FULL) return lane_step_neq<1>(v);
        else return lane_read_neq(v, sp1);
    };
    // ---- 3a. PCR levels with distance < R (partly inside the lane) -------
#pragma unroll
    for (int p = 0; (1 << p) < R; ++p) {
        const int dl = 1 << p;
        NEq<T> fromL[R], fromR[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (k < dl) {
                fromL[k] = left1(q[R - dl + k]);
                fromR[k] = right1(q[k]);
            }
        }
        NEq<T> nw[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const NEq<T> lft = (r - dl >= 0) ? q[(r - dl >= 0) ? r - dl : 0]
                                             : fromL[(r - dl >= 0) ? 0 : r];
            const NEq<T> rgt = (r + dl < R) ? q[(r + dl < R) ? r + dl : 0]
                                            : fromR[(r + dl < R) ? 0 : r + dl - R];
            nw[r] = reduce_neq(q[r], lft, rgt);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) q[r] = nw[r];
    }
    // ---- 3b. PCR levels with distance >= R: lane distance 1 .. L/4 -------
    if (1 < (L >> 1)) {
#pragma unroll
        for (int r = 0; r < R; ++r) q[r] = reduce_neq(q[r], left1(q[r]), right1(q[r]));
    }
    // (row r of a lane meets only row r of other lanes here, so each is
    // updated in place once its neighbours are read)
    for (int ld = 2; ld < (L >> 1); ld <<= 1) {
        const int sl = src(-ld), sr = src(ld);
#pragma unroll
        for (int r = 0; r < R; ++r)
            q[r] = reduce_neq(q[r], lane_read_neq(q[r], sl), lane_read_neq(q[r], sr));
    }
    // ---- 4. 2x2 solves at lane distance L/2 ------------------------------
    // Row i of the lower half has a = 0 exactly and its partner (upper half)
    // c = 0, and the other way round, so with k = a + c of the own row and
    // m = a + c of the partner row both lanes solve
    //   x = (d - k d_partner) / (1 - k m)
    // (x1 = (d1 - c1 d2) / (1 - c1 a2), x2 = (d2 - a2 d1) / (1 - a2 c1)).
    const int half = L >> 1;
    const bool lower = ll < half;
    auto solve_pairs = [&](auto partner_of) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const NEq<T> p = partner_of(q[r]);
            const T k = q[r].a + q[r].c, m = p.a + p.c;
            xv[S * r + S - 1] = (q[r].d - k * p.d) * tdiv(T(1), T(1) - k * m);
        }
    };
    if (L == 64)
        solve_pairs([&](const NEq<T>& v) { return lane_xor_neq<32>(v, lane); });
    else if (L == 32)
        solve_pairs([&](const NEq<T>& v) { return lane_xor_neq<16>(v, lane); });
    else
        solve_pairs([&](const NEq<T>& v) {
            return lane_read_neq(v, lower ? lane + half : lane - half);
        });
}

template <typename T, int M, int R, bool FULL = false>
__device__ __forceinline__ void wave_core(Eq<T> (&e)[M], T (&xv)[M], int lane, int ll, int L,
                                          int rest)
{
    constexpr int S = M / R;             // in-lane stride of the PCR equations
    constexpr int KIN = Log2<S>::value;  // in-lane CR levels

    // ---- 1. in-lane CR forward levels -----------------------------------
#pragma unroll
    for (int k = 1; k <= KIN; ++k) {
        const int SS = 1 << k, H = SS >> 1;
        // The top equation of this lane needs the right lane's equation at
        // slot H-1, which that lane finished at level k-1.
        const Eq<T> nb = lane_step_eq<1>(e[H - 1], ll + 1 < L);
#pragma unroll
        for (int j = 0; j < M / SS; ++j) {
            const int i = SS * j + SS - 1;
            const int ir = (i + H < M) ? i + H : i;
            const Eq<T> rgt = (i + H < M) ? e[ir] : nb;
            e[i] = reduce_eq(e[i], e[i - H], rgt);
        }
    }

#pragma unroll
    for (int s = 0; s < M; ++s) xv[s] = T(0);

    if constexpr (R == 1) {
        // ---- 2. cross-lane CR levels on slot M-1 while count > rest ------
        // (the first level reads the neighbouring lanes by DPP, the later
        // ones by ds_bpermute; the same for steps 3 and 5a)
        int hs = 1;
        int count = L;
        auto cr_level = [&](const Eq<T>& lft, const Eq<T>& rgt) {
            const bool active = ((ll + 1) & (2 * hs - 1)) == 0;
            const Eq<T> r = reduce_eq(e[M - 1], lft, rgt);
            if (active) e[M - 1] = r;
            count >>= 1;
            hs <<= 1;
        };
        if (count > rest)
            cr_level(lane_step_eq<-1>(e[M - 1], ll >= 1), lane_step_eq<1>(e[M - 1], ll + 1 < L));
        while (count > rest)
            cr_level(lane_read_eq(e[M - 1], lane - hs, ll - hs >= 0),
                     lane_read_eq(e[M - 1], lane + hs, ll + hs < L));
        // hs = L / rest: spacing of the lanes that hold the rest system
        const int LS = hs;
        const bool pact = ((ll + 1) & (LS - 1)) == 0;
        // ---- 3. PCR: log2(rest/2) levels, lane distance LS .. L/4 ----------
        int dist = LS;
        auto pcr_level = [&](const Eq<T>& lft, const Eq<T>& rgt) {
            const Eq<T> r = reduce_eq(e[M - 1], lft, rgt);
            if (pact) e[M - 1] = r;
            dist <<= 1;
        };
        if (dist == 1 && dist < (L >> 1))
            pcr_level(lane_step_eq<-1>(e[M - 1], ll >= 1), lane_step_eq<1>(e[M - 1], ll + 1 < L));
        while (dist < (L >> 1))
            pcr_level(lane_read_eq(e[M - 1], lane - dist, ll - dist >= 0),
                      lane_read_eq(e[M - 1], lane + dist, ll + dist < L));
        // ---- 4. 2x2 solves at lane distance L/2 --------------------------
        const int half = L >> 1;
        const bool lower = ll < half;
        Eq<T> partner;
        if (L == 64) partner = lane_xor_eq<32>(e[M - 1], lane);
        else if (L == 32) partner = lane_xor_eq<16>(e[M - 1], lane);
        else partner = lane_read_eq(e[M - 1], lower ? lane + half : lane - half, true);
        {
            const Eq<T> lo = lower ? e[M - 1] : partner;
            const Eq<T> hi = lower ? partner : e[M - 1];
            T x1, x2;
            solve2(lo, hi, x1, x2);
            xv[M - 1] = lower ? x1 : x2;
        }
        // ---- 5a. cross-lane CR back-substitution -------------------------
        auto bs_level = [&](int h, T xl, T xr) {
            const bool solve = ((ll + 1) & (2 * h - 1)) == h;
            const T v = backsub(e[M - 1], xl, xr);
            if (solve) xv[M - 1] = v;
        };
        for (int h = LS >> 1; h >= 2; h >>= 1)
            bs_level(h, lane_read_or_zero(xv[M - 1], lane - h, ll - h >= 0),
                     lane_read_or_zero(xv[M - 1], lane + h, ll + h < L));
        if (LS >= 2)
            bs_level(1, lane_step_or_zero<-1>(xv[M - 1], ll >= 1),
                     lane_step_or_zero<1>(xv[M - 1], ll + 1 < L));
    } else {
        pcr_rows<T, M, R, FULL>(e, xv, lane, ll, L);
    }

    // ---- 5b. in-lane CR back-substitution --------------------------------
    const T xprev = lane_step_or_zero<-1>(xv[M - 1], ll >= 1);
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

}

// Bytes of the reduced-system area per lane of the workgroup: W wave-top
// equations, or two buffers of up to 64 W unit-diagonal rows for the PCR
// levels across the workgroup.
template <typename T>
constexpr size_t wg_red_bytes_per_lane()
{
    return sizeof(Eq<T>) > 2 * sizeof(NEq<T>) ? sizeof(Eq<T>) : 2 * sizeof(NEq<T>);
}

template <typename T, int M>
constexpr size_t wg_core_lds_bytes(int W)
{
    return (size_t)W * (Log2<M>::value + 6) * sizeof(Eq<T>) +
           (size_t)kWave * W * wg_red_bytes_per_lane<T>() + (size_t)kWave * W * sizeof(T);
}

// Workgroup barrier of the multi-wave core. DMA: the caller has global->LDS
// copies in flight (global_load_lds into an LDS area the core never
// touches). __syncthreads() would wait for them: its release fence covers
// LDS writes, and those copies are LDS writes counted by vmcnt. This form
// waits for the wave's own LDS operations only (ds_* and ds_bpermute, all
// counted by lgkmcnt) and then meets the other waves; the "memory" clobber
// keeps the compiler from moving memory accesses across it.
template <bool DMA>
__device__ __forceinline__ void wg_barrier()
{
    if constexpr (DMA) asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    else __syncthreads();
}

// s_waitcnt immediate (gfx9 encoding) for vmcnt(0) alone: vmcnt bits 3:0
// and 15:14 zero, expcnt (6:4) and lgkmcnt (11:8) at their maxima. Given to
// __builtin_amdgcn_s_waitcnt, the wait is one the compiler's own wait
// insertion sees (global->LDS copies it then knows to have landed).
constexpr int kWaitVmcnt0 = 0x0F70;

// Multi-wave form: W waves of one workgroup hold one system, wave w owning
// unknowns [w*64*M, (w+1)*64*M) (see kern_reg_mw.hip). Every thread of the
// workgroup must call it (it contains barriers). `xchg` is LDS of
// wg_core_lds_bytes(W): W*(log2(M)+6) prefix equations, the reduced-system
// area and 64*W values; the caller's staging area may alias it (a barrier is
// taken before first use). DMA: see wg_barrier.
template <typename T, int M, bool DMA = false>
__device__ __forceinline__ void wg_core(Eq<T> (&e)[M], T (&xv)[M], int w, int W, int lane,
                                        int rest, char* xchg)
{
    constexpr int LOGM = Log2<M>::value;
    constexpr int NPRE = LOGM + 6;  // published prefix equations per wave
    Eq<T>* pre = reinterpret_cast<Eq<T>*>(xchg);   // [W][NPRE]
    Eq<T>* red = pre + W * NPRE;                     // [W] wave tops, or the PCR rows
    T* xs = reinterpret_cast<T*>(reinterpret_cast<char*>(red) +
                                 (size_t)kWave * W * wg_red_bytes_per_lane<T>());  // [64*W]

    const Eq<T> top0 = e[M - 1];  // level-0 top equation (lane 63's chain restarts here)
    wg_barrier<DMA>();              // staging rows become the exchange area
    if constexpr (kLdsPoison) {
        // the W waves of this system fill its exchange area (other groups of
        // the workgroup fill theirs; every thread reaches the barrier)
        lds_fill_nan(xchg, (unsigned)wg_core_lds_bytes<T, M>(W), w * kWave + lane, W * kWave);
        wg_barrier<DMA>();
    }

    // ---- in-lane CR levels (lane 63's top slot is provisional) ------------
#pragma unroll
    for (int k = 1; k <= LOGM; ++k) {
        const int SS = 1 << k, H = SS >> 1;
        const Eq<T> nb = lane_step_eq<1>(e[H - 1], lane + 1 < kWave);
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
    wg_barrier<DMA>();

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

#pragma unroll
    for (int s = 0; s < M; ++s) xv[s] = T(0);
    T xprevw;  // top unknown of the previous wave (0 for wave 0)

    if (rest <= W) {
        // ---- wave-top system: W equations, solved by wave 0 ---------------
        if (lane == kWave - 1) red[w] = e[M - 1];
        wg_barrier<DMA>();
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
        wg_barrier<DMA>();
        if (lane == kWave - 1) xv[M - 1] = xs[w];
        xprevw = w > 0 ? xs[w - 1] : T(0);
    } else {
        // ---- PCR on `rest` equations spread over the workgroup ------------
        // Rows scaled to a unit diagonal (three values per row, one
        // reciprocal per reduction); level k reads buffer k&1 and writes the
        // other, so one barrier per level suffices. A neighbour outside the
        // system is the zero row: it meets a coupling that is exactly 0.
        const int LS = hs;  // lanes between consecutive PCR equations
        const bool pact = ((lane + 1) & (LS - 1)) == 0;
        const int q = (w * kWave + lane + 1) / LS - 1;
        // two buffers of `rest` rows, each as (a, c) pairs then d values
        // (aligned 8- / 16-byte and 4- / 8-byte LDS accesses)
        using T2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
        T* const rows = reinterpret_cast<T*>(red);
        auto put = [&](int b, int i, const NEq<T>& v) {
            T* bb = rows + (size_t)b * 3 * rest;
            reinterpret_cast<T2*>(bb)[i] = T2{v.a, v.c};
            bb[2 * rest + i] = v.d;
        };
        auto get = [&](int b, int i) {
            const T* bb = rows + (size_t)b * 3 * rest;
            const T2 ac = reinterpret_cast<const T2*>(bb)[i];
            return NEq<T>{ac.x, ac.y, bb[2 * rest + i]};
        };
        NEq<T> me = NEq<T>{T(0), T(0), T(0)};
        if (pact) {
            me = normalize_eq(e[M - 1]);
            put(0, q, me);
        }
        wg_barrier<DMA>();
        int cb = 0;  // buffer holding the current rows
        for (int delta = 1; delta < (rest >> 1); delta <<= 1) {
            if (pact) {
                const NEq<T> zero{T(0), T(0), T(0)};
                const NEq<T> lft = q - delta >= 0 ? get(cb, q - delta) : zero;
                const NEq<T> rgt = q + delta < rest ? get(cb, q + delta) : zero;
                me = reduce_neq(me, lft, rgt);
                put(cb ^ 1, q, me);
            }
            wg_barrier<DMA>();
            cb ^= 1;
        }
        if (pact) {
            // x_q + c x_{q+half} = d (lower half), a x_{q-half} + x_q = d (upper)
            const int half = rest >> 1;
            const bool lower = q < half;
            const NEq<T> p = get(cb, lower ? q + half : q - half);
            const T cm = lower ? me.c : me.a;  // my coupling to the partner
            const T cp = lower ? p.a : p.c;    // the partner's coupling to me
            xv[M - 1] = tdiv(me.d - cm * p.d, T(1) - cm * cp);
            xs[q] = xv[M - 1];
        }
        wg_barrier<DMA>();
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

}


}  // namespace trid
