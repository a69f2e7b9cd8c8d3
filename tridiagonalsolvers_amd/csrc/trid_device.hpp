// trid_device.hpp -- device-side building blocks shared by the gfx950 kernels.
//
// One tridiagonal equation i reads  a_i x_{i-1} + b_i x_i + c_i x_{i+1} = d_i.
// After reduction levels it couples x_i to x_{i-delta} and x_{i+delta} with
// the same four coefficients. A neighbour that lies outside the system is an
// identity row (a = c = d = 0, b = 1): it contributes nothing. This is what
// the reference obtains by clamping neighbour indices and relying on
// a[0] = c[n-1] = 0 (cr_kernel.cu:52, pcr_kernel.cu:33-37); an explicit
// identity row also makes padding to a power of two exact.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace trid {

constexpr int kWave = 64;

template <typename T>
struct Eq {
    T a, b, c, d;
};

template <typename T>
__device__ __forceinline__ Eq<T> identity_eq()
{
    return Eq<T>{T(0), T(1), T(0), T(0)};
}

// Quotient p / q. FP32 uses the hardware reciprocal (1 ulp) and one
// multiply. FP64 refines the hardware reciprocal (about 23 correct bits) by
// two Newton-Raphson steps, r <- r + r (1 - q r), each of which doubles the
// correct bits, then multiplies: within about one ulp of p / q for finite,
// nonzero q, in a third of the instructions of the IEEE-rounded division
// sequence (on 4,096 systems of 8,192 unknowns the exact division costs 14%
// of the FP64 solve). A zero q gives a non-finite quotient, as division does.
__device__ __forceinline__ float tdiv(float p, float q)
{
    return p * __builtin_amdgcn_rcpf(q);
}
__device__ __forceinline__ double tdiv(double p, double q)
{
    double r = __builtin_amdgcn_rcp(q);
    r = __builtin_fma(__builtin_fma(-q, r, 1.0), r, r);
    r = __builtin_fma(__builtin_fma(-q, r, 1.0), r, r);
    return p * r;
}

// One CR/PCR reduction step (cr_kernel.cu:53-58, pcr_kernel.cu:39-56):
//   k1 = a_i / b_L, k2 = c_i / b_R
//   b_i <- b_i - c_L k1 - a_R k2,  d_i <- d_i - d_L k1 - d_R k2
//   a_i <- -a_L k1,                c_i <- -c_R k2
template <typename T>
__device__ __forceinline__ Eq<T> reduce_eq(const Eq<T>& e, const Eq<T>& L, const Eq<T>& R)
{
    T k1 = tdiv(e.a, L.b);
    T k2 = tdiv(e.c, R.b);
    Eq<T> r;
    r.b = e.b - L.c * k1 - R.a * k2;
    r.d = e.d - L.d * k1 - R.d * k2;
    r.a = -L.a * k1;
    r.c = -R.c * k2;
    return r;
}

// 2x2 solve by Cramer's rule (cr_kernel.cu:65-69) for the pair
// (e1 at the lower index, e2 at the upper index).
template <typename T>
__device__ __forceinline__ void solve2(const Eq<T>& e1, const Eq<T>& e2, T& x1, T& x2)
{
    T det = e2.b * e1.b - e1.c * e2.a;
    x1 = tdiv(e2.b * e1.d - e1.c * e2.d, det);
    x2 = tdiv(e2.d * e1.b - e1.d * e2.a, det);
}

// Back-substitution of one unknown (cr_kernel.cu:80-84):
//   x_i = (d_i - a_i x_{i-delta} - c_i x_{i+delta}) / b_i
template <typename T>
__device__ __forceinline__ T backsub(const Eq<T>& e, T xl, T xr)
{
    return tdiv(e.d - e.a * xl - e.c * xr, e.b);
}

// Row scaled to a unit diagonal: x_i + a x_{i-delta} + c x_{i+delta} = d.
// PCR on such rows moves three values per row between lanes instead of four
// and takes one reciprocal per reduction instead of two.
template <typename T>
struct NEq {
    T a, c, d;
};

template <typename T>
__device__ __forceinline__ NEq<T> normalize_eq(const Eq<T>& e)
{
    const T r = tdiv(T(1), e.b);
    return NEq<T>{e.a * r, e.c * r, e.d * r};
}

// PCR reduction of unit-diagonal rows (the step of reduce_eq, pcr_kernel.cu:39-56,
// divided through by the new diagonal):
//   beta = 1 - a_i c_L - c_i a_R
//   a_i <- -a_i a_L / beta,  c_i <- -c_i c_R / beta
//   d_i <- (d_i - a_i d_L - c_i d_R) / beta
// A neighbour lying outside the system meets a coefficient that is exactly 0
// (a_i = 0 while i - delta < 0, c_i = 0 while i + delta >= n, kept so by the
// step itself), so any finite row may stand in for it.
template <typename T>
__device__ __forceinline__ NEq<T> reduce_neq(const NEq<T>& e, const NEq<T>& L, const NEq<T>& R)
{
    const T r = tdiv(T(1), T(1) - e.a * L.c - e.c * R.a);
    const T ar = e.a * r, cr = e.c * r;
    return NEq<T>{-ar * L.a, -cr * R.c, e.d * r - ar * L.d - cr * R.d};
}

// ---------------------------------------------------------------------------
// cross-lane moves (ds_bpermute through the LDS crossbar; no LDS storage)
// ---------------------------------------------------------------------------

__device__ __forceinline__ float lane_read(float v, int src)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
}
__device__ __forceinline__ double lane_read(double v, int src)
{
    int lo = __builtin_amdgcn_ds_bpermute(src << 2, __double2loint(v));
    int hi = __builtin_amdgcn_ds_bpermute(src << 2, __double2hiint(v));
    return __hiloint2double(hi, lo);
}

// Equation held by lane `src` (all lanes must execute this), replaced by the
// identity row when `valid` is false for the reading lane.
template <typename T>
__device__ __forceinline__ Eq<T> lane_read_eq(const Eq<T>& e, int src, bool valid)
{
    src &= (kWave - 1);
    Eq<T> r;
    r.a = lane_read(e.a, src);
    r.b = lane_read(e.b, src);
    r.c = lane_read(e.c, src);
    r.d = lane_read(e.d, src);
    if (!valid) r = identity_eq<T>();
    return r;
}

template <typename T>
__device__ __forceinline__ T lane_read_or_zero(T v, int src, bool valid)
{
    T r = lane_read(v, src & (kWave - 1));
    return valid ? r : T(0);
}

// Value of lane (lane + D), D = +1 or -1, by a DPP wave shift (wave_shl:1 /
// wave_shr:1): a VALU move that reads the neighbouring lane's register, with
// no trip through the LDS crossbar that ds_bpermute takes. The lane shifted in
// at the end of the wave reads 0 (bound_ctrl set: an out-of-range source
// reads 0, so the move needs no prior copy of an old value into its
// destination, one VALU instruction fewer per move); callers mask the lanes
// whose neighbour lies outside their system, as they do for ds_bpermute reads. D is a template
// argument so that the choice of instruction is made at compile time. Call
// it where every lane of the wave is active (not in a branch that some lanes
// skip): a lane whose source lane is inactive reads 0.
template <int D>
__device__ __forceinline__ int lane_step_i32(int v)
{
    static_assert(D == 1 || D == -1, "DPP wave shift by one lane");
    return __builtin_amdgcn_update_dpp(0, v, D == 1 ? 0x130 : 0x138, 0xf, 0xf, true);
}
template <int D>
__device__ __forceinline__ float lane_step(float v)
{
    return __int_as_float(lane_step_i32<D>(__float_as_int(v)));
}
template <int D>
__device__ __forceinline__ double lane_step(double v)
{
    const int lo = lane_step_i32<D>(__double2loint(v));
    const int hi = lane_step_i32<D>(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

// Equation of lane (lane + D); the identity row where `valid` is false.
template <int D, typename T>
__device__ __forceinline__ Eq<T> lane_step_eq(const Eq<T>& e, bool valid)
{
    Eq<T> r;
    r.a = lane_step<D>(e.a);
    r.b = lane_step<D>(e.b);
    r.c = lane_step<D>(e.c);
    r.d = lane_step<D>(e.d);
    if (!valid) r = identity_eq<T>();
    return r;
}

template <int D, typename T>
__device__ __forceinline__ T lane_step_or_zero(T v, bool valid)
{
    const T r = lane_step<D>(v);
    return valid ? r : T(0);
}

// Value of lane (lane ^ X), X = 32 or 16, by gfx950's v_permlane32_swap /
// v_permlane16_swap: the swap of a register with itself leaves, in its two
// results, each half (row pair) of the wave next to its partner half, and the
// lane picks the one that is not its own. A VALU operation, no LDS crossbar.
template <int X>
__device__ __forceinline__ int lane_xor_i32(int v, int lane)
{
    static_assert(X == 32 || X == 16, "v_permlane32_swap / v_permlane16_swap");
    if constexpr (X == 32) {
        const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return (lane & 32) ? r[0] : r[1];
    } else {
        const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return (lane & 16) ? r[0] : r[1];
    }
}
template <int X>
__device__ __forceinline__ float lane_xor(float v, int lane)
{
    return __int_as_float(lane_xor_i32<X>(__float_as_int(v), lane));
}
template <int X>
__device__ __forceinline__ double lane_xor(double v, int lane)
{
    const int lo = lane_xor_i32<X>(__double2loint(v), lane);
    const int hi = lane_xor_i32<X>(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}
template <int X, typename T>
__device__ __forceinline__ Eq<T> lane_xor_eq(const Eq<T>& e, int lane)
{
    return Eq<T>{lane_xor<X>(e.a, lane), lane_xor<X>(e.b, lane), lane_xor<X>(e.c, lane),
                 lane_xor<X>(e.d, lane)};
}

// The same moves for unit-diagonal rows (three values each).
template <typename T>
__device__ __forceinline__ NEq<T> lane_read_neq(const NEq<T>& e, int src)
{
    src &= (kWave - 1);
    return NEq<T>{lane_read(e.a, src), lane_read(e.c, src), lane_read(e.d, src)};
}
template <int D, typename T>
__device__ __forceinline__ NEq<T> lane_step_neq(const NEq<T>& e)
{
    return NEq<T>{lane_step<D>(e.a), lane_step<D>(e.c), lane_step<D>(e.d)};
}
template <int X, typename T>
__device__ __forceinline__ NEq<T> lane_xor_neq(const NEq<T>& e, int lane)
{
    return NEq<T>{lane_xor<X>(e.a, lane), lane_xor<X>(e.c, lane), lane_xor<X>(e.d, lane)};
}

// Value of `v` in lane `src` (wave-uniform index), broadcast to all lanes.
__device__ __forceinline__ float lane_bcast(float v, int src)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ __forceinline__ double lane_bcast(double v, int src)
{
    int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}
template <typename T>
__device__ __forceinline__ Eq<T> lane_bcast_eq(const Eq<T>& e, int src)
{
    return Eq<T>{lane_bcast(e.a, src), lane_bcast(e.b, src), lane_bcast(e.c, src),
                 lane_bcast(e.d, src)};
}

// ---------------------------------------------------------------------------
// LDS poisoning (debug builds compiled with -DTRID_LDS_POISON): every LDS
// staging and exchange area is filled with all-ones bits -- a NaN in FP32 and
// in FP64 -- before each use, so that a read of an element that no thread
// wrote since turns into NaN in the results and fails the parity tests.
// Release builds compile all of this out.
// ---------------------------------------------------------------------------
#ifdef TRID_LDS_POISON
constexpr bool kLdsPoison = true;
#else
constexpr bool kLdsPoison = false;
#endif

// Threads t = 0..nthr-1 fill [p, p + bytes) (p 4-byte aligned, bytes a
// multiple of 4) with 4-byte stores.
__device__ __forceinline__ void lds_fill_nan(void* p, unsigned bytes, int t, int nthr)
{
    if constexpr (kLdsPoison) {
        unsigned* w = static_cast<unsigned*>(p);
        for (unsigned i = t; i < bytes / 4; i += nthr) w[i] = ~0u;
    }
}
// Area used by the whole workgroup; every thread calls it (barrier).
__device__ __forceinline__ void lds_poison_wg(void* p, unsigned bytes)
{
    if constexpr (kLdsPoison) {
        lds_fill_nan(p, bytes, (int)threadIdx.x, (int)blockDim.x);
        __syncthreads();
    }
}
// Area of one wave; every lane of the wave calls it. LDS operations of a
// wave complete in order, so the wave's later accesses see the fill.
__device__ __forceinline__ void lds_poison_wave(void* p, unsigned bytes, int lane)
{
    if constexpr (kLdsPoison) {
        lds_fill_nan(p, bytes, lane, kWave);
        __builtin_amdgcn_wave_barrier();
    }
}

// Workgroup b of a grid padded to a multiple of 8 Q -> the column group it
// takes, for kernels whose consecutive column groups are contiguous bytes of
// every row (interleaved batches). The 8 XCDs take the workgroups in turn,
// so b, b + 8, .., b + 8 (Q-1) run at about the same time on one XCD: they
// take Q consecutive groups, and HBM sees their rows as whole spans of Q
// groups instead of pieces of one. (Measured neutral for the lane-per-system
// pivoting and elimination kernels, which keep the plain order.)
template <int Q>
__device__ __forceinline__ int64_t xcd_span_group(int64_t b)
{
    const int64_t x = b % 8, k = b / 8;
    return Q * ((k / Q) * 8 + x) + (k % Q);
}
// Groups of `group_bytes` bytes per row that make a span of `span_bytes`.
constexpr int xcd_span(int span_bytes, int group_bytes)
{
    return span_bytes / group_bytes > 1 ? span_bytes / group_bytes : 1;
}

template <int V>
struct Log2 {
    static constexpr int value = 1 + Log2<V / 2>::value;
};
template <>
struct Log2<1> {
    static constexpr int value = 0;
};

}  // namespace trid
