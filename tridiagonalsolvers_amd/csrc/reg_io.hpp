// reg_io.hpp -- global <-> register movement of the register-resident
// kernels (kern_reg.hip, kern_reg_mw.hip).
#pragma once

#include "trid_device.hpp"

namespace trid {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// I/O modes of the dense path (see RegArgs::io):
//   0  per-lane 16-byte accesses, non-temporal
//   1  per-lane 16-byte accesses, default cache policy
//   2  wave-coalesced 16-byte accesses staged through LDS (padded rows)
//   3  as 2, non-temporal
template <bool NT>
static __device__ __forceinline__ u32x4 gload16(const void* p)
{
    if constexpr (NT) return __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
    else return *reinterpret_cast<const u32x4*>(p);
}
template <bool NT>
static __device__ __forceinline__ void gstore16(void* p, u32x4 v)
{
    if constexpr (NT) __builtin_nontemporal_store(v, reinterpret_cast<u32x4*>(p));
    else *reinterpret_cast<u32x4*>(p) = v;
}

template <typename T, int M>
static __device__ __forceinline__ void set_identity(Eq<T> (&e)[M])
{
#pragma unroll
    for (int s = 0; s < M; ++s) e[s] = identity_eq<T>();
}

// Per-lane 16-byte loads of the lane's M consecutive unknowns.
// Constant coefficients (cc): a, b, c are the same scalars for every row
// and only d is read.
template <typename T>
struct ConstCoef {
    bool on;
    T a, b, c;
};
template <typename T>
constexpr ConstCoef<T> kNoConstCoef{false, T(0), T(0), T(0)};

template <typename T, int M, bool NT>
static __device__ __forceinline__ void load_direct(const T* __restrict__ A, const T* __restrict__ B,
                                            const T* __restrict__ C, const T* __restrict__ D,
                                            int64_t off, Eq<T> (&e)[M], ConstCoef<T> cc)
{
    constexpr int E = 16 / sizeof(T);
    T ta[M], tb[M], tc[M], td[M];
#pragma unroll
    for (int q = 0; q < M / E; ++q) {
        u32x4 vd = gload16<NT>(D + off + q * E);
        __builtin_memcpy(&td[q * E], &vd, 16);
        if (!cc.on) {
            u32x4 va = gload16<NT>(A + off + q * E);
            u32x4 vb = gload16<NT>(B + off + q * E);
            u32x4 vc = gload16<NT>(C + off + q * E);
            __builtin_memcpy(&ta[q * E], &va, 16);
            __builtin_memcpy(&tb[q * E], &vb, 16);
            __builtin_memcpy(&tc[q * E], &vc, 16);
        }
    }
#pragma unroll
    for (int s = 0; s < M; ++s)
        e[s] = cc.on ? Eq<T>{cc.a, cc.b, cc.c, td[s]} : Eq<T>{ta[s], tb[s], tc[s], td[s]};
}

template <typename T, int M, bool NT>
static __device__ __forceinline__ void store_direct(T* __restrict__ X, int64_t off, const T (&xv)[M])
{
    constexpr int E = 16 / sizeof(T);
#pragma unroll
    for (int q = 0; q < M / E; ++q) {
        u32x4 v;
        __builtin_memcpy(&v, &xv[q * E], 16);
        gstore16<NT>(X + off + q * E, v);
    }
}

// LDS row of one lane: M elements plus 16 bytes of padding. Unpadded rows
// (M*sizeof(T) in {32, 64, 128} bytes) would put the row-wise 16-byte reads
// of neighbouring lanes on the same banks; the extra 16 bytes shift each
// row by four banks. (Measured on 65,536 x 1,024 FP32: bank-conflict cycles
// are a quarter of its LDS-busy cycles, and its time is set by HBM.)
template <typename T, int M>
struct Stage {
    static constexpr int E = 16 / sizeof(T);
    static constexpr int Q = M / E;                        // 16-byte chunks per row
    static constexpr int kRowBytes = M * (int)sizeof(T) + 16;
    static constexpr int kWaveBytes = kWave * kRowBytes;
    // chunk k of the wave block (k = 64 q + lane) lives in row k / Q
    __device__ static __forceinline__ int chunk_addr(int k)
    {
        return (k / Q) * kRowBytes + (k % Q) * 16;
    }
};

// Element offset of position o of a wave's block of 64*M unknowns. The block
// is contiguous (seg_log2 < 0), or (strided batches) made of systems of
// 2^seg_log2 unknowns that start `seg_stride` elements apart.
static __device__ __forceinline__ int64_t block_elem(int64_t wave_base, int64_t o, int seg_log2,
                                                     int64_t seg_stride)
{
    if (seg_log2 < 0) return wave_base + o;
    return wave_base + (o >> seg_log2) * seg_stride + (o & ((int64_t(1) << seg_log2) - 1));
}

// Wave-coalesced loads: instruction q of the wave reads the q-th contiguous
// 1 KiB of the wave's block (for a strided batch: 16-byte pieces of its
// systems); the chunks are transposed through LDS so that lane l ends up with
// unknowns [l*M, l*M+M) of the block.
template <typename T, int M, bool NT>
static __device__ __forceinline__ void load_staged(const T* __restrict__ A, const T* __restrict__ B,
                                            const T* __restrict__ C, const T* __restrict__ D,
                                            int64_t wave_base, int64_t total, int lane,
                                            char* lds, Eq<T> (&e)[M], ConstCoef<T> cc,
                                            int seg_log2 = -1, int64_t seg_stride = 0)
{
    using S = Stage<T, M>;
    constexpr int E = S::E, Q = S::Q;
    u32x4 r[4][Q];
    const T* src[4] = {A, B, C, D};
    const int first = cc.on ? 3 : 0;  // constant coefficients: d only
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int64_t el =
            block_elem(wave_base, (int64_t)(kWave * q + lane) * E, seg_log2, seg_stride);
        const bool ok = el < total;
#pragma unroll
        for (int arr = 0; arr < 4; ++arr)
            r[arr][q] = (ok && arr >= first) ? gload16<NT>(src[arr] + el) : u32x4{0u, 0u, 0u, 0u};
    }
    T v[4][M];
    if (cc.on) {
#pragma unroll
        for (int s = 0; s < M; ++s) {
            v[0][s] = cc.a;
            v[1][s] = cc.b;
            v[2][s] = cc.c;
        }
    }
#pragma unroll
    for (int arr = 0; arr < 4; ++arr) {
        if (arr < first) continue;
        lds_poison_wave(lds, S::kWaveBytes, lane);
#pragma unroll
        for (int q = 0; q < Q; ++q)
            *reinterpret_cast<u32x4*>(lds + S::chunk_addr(kWave * q + lane)) = r[arr][q];
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            u32x4 t = *reinterpret_cast<const u32x4*>(lds + lane * S::kRowBytes + q * 16);
            __builtin_memcpy(&v[arr][q * E], &t, 16);
        }
        __builtin_amdgcn_wave_barrier();
    }
#pragma unroll
    for (int s = 0; s < M; ++s) e[s] = Eq<T>{v[0][s], v[1][s], v[2][s], v[3][s]};
}

template <typename T, int M, bool NT>
static __device__ __forceinline__ void store_staged(T* __restrict__ X, int64_t wave_base, int64_t total,
                                             int lane, char* lds, const T (&xv)[M],
                                             int seg_log2 = -1, int64_t seg_stride = 0)
{
    using S = Stage<T, M>;
    constexpr int E = S::E, Q = S::Q;
    lds_poison_wave(lds, S::kWaveBytes, lane);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        u32x4 t;
        __builtin_memcpy(&t, &xv[q * E], 16);
        *reinterpret_cast<u32x4*>(lds + lane * S::kRowBytes + q * 16) = t;
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const u32x4 t = *reinterpret_cast<const u32x4*>(lds + S::chunk_addr(kWave * q + lane));
        const int64_t el =
            block_elem(wave_base, (int64_t)(kWave * q + lane) * E, seg_log2, seg_stride);
        if (el < total) gstore16<NT>(X + el, t);
    }
}


}  // namespace trid
