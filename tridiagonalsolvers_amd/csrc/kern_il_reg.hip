// kern_il_reg.hip -- interleaved batches (row i of system s at i*stride + s)
// solved by the register-resident CR-PCR core instead of a lane-per-system
// elimination.
//
// No counterpart in the reference (system-major only). A lane-per-system
// elimination reads every row coalesced across systems but must keep its
// factors in memory: 9 elements of traffic per unknown. Here a workgroup of
// S waves takes S consecutive systems (S = 16 FP32, 8 FP64) and gives each
// wave one system in the layout of the register core (lane l holds unknowns
// [MI l, MI l + MI), MI = 16, 8 or 4 unknowns per lane for n up to 1024, 512
// or 256), so the solve moves only the 5 elements per unknown it must (read
// a, b, c, d; write x). The gather is the transposition: a row of the S
// systems is 64 contiguous bytes, read as 16-byte vectors, one per thread and
// array per chunk of rows; the chunk is scattered into LDS by system and lane
// group, and the lanes of every wave that own those rows read their MI
// consecutive rows back as 16-byte vectors. The solution leaves the same way.
// LDS layout per array: [system][lane group][MI rows + 16 bytes of padding],
// plus 16 bytes between systems: the per-lane vector reads fall on distinct
// banks and the scattering writes of a half-wave conflict at most 2-way.
//
// Rows past n are identity rows and systems past N are never stored, so any
// n in the range and any N work; a[first] and c[last] are ignored. x may be
// d: a workgroup reads all of its systems' rows before it writes any x.
#include "reg_core.hpp"
#include "reg_io.hpp"
#include "trid_device.hpp"
#include "trid_launch.hpp"

namespace trid {

namespace {

// S systems (waves) per workgroup (TRID_IL_SYSTEMS_F32 in FP32, 8 in FP64:
// the FP64 core needs about 210 VGPRs). A row of the workgroup's systems is
// S * sizeof(T) bytes, a part of a 128-byte line shared by P workgroups.
// FP32 at 16 (1,024 threads, one workgroup per CU) takes 444-463 us on
// 65,536 x 1,024; at 8 (512 threads, two workgroups per CU, 32-byte row
// parts, P = 4) 564 us: the narrower parts cost more than the overlap of
// two workgroups gains.
#ifndef TRID_IL_SYSTEMS_F32
#define TRID_IL_SYSTEMS_F32 16
#endif
// ... and 8 at 8 or 4 unknowns per lane, where 16 systems' LDS (99 / 132 KB)
// leaves one workgroup per CU and 8 systems' (50 / 66 KB) two or three
// (131,072 x 512: 348 us against 379-386 us; at 16 per lane 8 systems take
// 419-421 us against 381-387 us, and FP64 at 4 systems 769-809 us against
// 661-772 us)
#ifndef TRID_IL_SYSTEMS_F32_SHORT
#define TRID_IL_SYSTEMS_F32_SHORT 8
#endif
#ifndef TRID_IL_SYSTEMS_F64
#define TRID_IL_SYSTEMS_F64 8
#endif
// Contiguous bytes of every row that the workgroups on one XCD take together
// (xcd_span_group).
constexpr int kIlSpanBytes = 1024;

template <typename T, int MI>
struct IlGeom {
    static constexpr int S = sizeof(T) == 8 ? TRID_IL_SYSTEMS_F64
                             : MI >= 16     ? TRID_IL_SYSTEMS_F32
                                            : TRID_IL_SYSTEMS_F32_SHORT;
    static constexpr int EV = 16 / (int)sizeof(T);           // elements per vector
    static constexpr int VPR = S / EV;                        // vectors per row
    static constexpr int RC = S * kWave / VPR;                // rows per chunk
    static constexpr int LC = RC / MI;                        // lanes per chunk
    static constexpr int NC = kWave / LC;                     // chunks per system
    static constexpr int GP = MI + EV;                        // padded lane group
    static constexpr int SP = LC * GP + EV;                   // system pitch
    static constexpr int AP = S * SP;                         // array pitch
    static constexpr int P = 128 / (S * (int)sizeof(T));      // workgroups per 128-byte line
    static constexpr int Q = xcd_span(kIlSpanBytes, S * (int)sizeof(T));  // workgroups per span
    static_assert(LC * NC == kWave && P * S * (int)sizeof(T) == 128, "chunk geometry");
};

// Workgroup -> group of S systems. A row of the group is 1/P of a 128-byte
// line; the other parts belong to the neighbouring groups. Workgroups are
// dealt to the 8 XCDs round-robin, so b, b + 8, .., b + 8 (Q-1) run at about
// the same time on one XCD: they take Q consecutive groups, i.e. the same
// Q * S * sizeof(T) contiguous bytes of every row. Every read of a line after
// the first hits that XCD's L2, and HBM sees whole 1 KiB pieces of each row
// (rows of an interleaved batch lie stride * sizeof(T) apart, a DRAM page
// or more). Measured on 65,536 x 1,024 (FP32 / FP64): Q = P (one 128-byte
// line): 449 / 937 us; 256 B: 437 / 892 us; 512 B: 403 / 796 us; 1 KiB:
// 381 / 770 us; 2 KiB: 395 / 798 us. (The grid is a multiple of 8 Q so the
// map is one-to-one: xcd_span_group.)

template <typename T, int MI>
__device__ __forceinline__ int il_addr(int sys, int row_in_chunk)
{
    using G = IlGeom<T, MI>;
    return sys * G::SP + (row_in_chunk / MI) * G::GP + (row_in_chunk % MI);
}

// One vector of array `arr` for chunk `c`: row c*RC + t/VPR, systems
// s0 + (t%VPR)*EV .. +EV-1. Identity values past n; nothing read past N.
template <typename T, int MI>
__device__ __forceinline__ void il_fetch(const T* __restrict__ P, int arr, int c, int t,
                                         int64_t n, int64_t N, int64_t stride, int64_t s0,
                                         bool vec, T (&v)[IlGeom<T, MI>::EV])
{
    using G = IlGeom<T, MI>;
    const int64_t row = (int64_t)c * G::RC + t / G::VPR;
    const int64_t s = s0 + (t % G::VPR) * G::EV;
    const T fill = arr == 1 ? T(1) : T(0);
    const int64_t off = row * stride + s;
    if (row < n && vec && s + G::EV <= N) {
        const u32x4 q = gload16<false>(P + off);
        __builtin_memcpy(v, &q, 16);
    } else {
#pragma unroll
        for (int e = 0; e < G::EV; ++e) v[e] = (row < n && s + e < N) ? P[off + e] : fill;
    }
    if (arr == 0 && row == 0) {
#pragma unroll
        for (int e = 0; e < G::EV; ++e) v[e] = T(0);
    }
    if (arr == 2 && row == n - 1) {
#pragma unroll
        for (int e = 0; e < G::EV; ++e) v[e] = T(0);
    }
}

template <typename T, int MI>
__global__ __launch_bounds__((IlGeom<T, MI>::S * kWave)) void il_reg_kernel(
    const T* __restrict__ A, const T* __restrict__ B, const T* __restrict__ C,
    const T* __restrict__ D, T* X, int64_t n, int64_t N, int64_t stride, int rest, int vec)
{
    using G = IlGeom<T, MI>;
    __shared__ __align__(16) T tile[4 * G::AP];
    const int t = threadIdx.x;
    const int w = t / kWave;
    const int lane = t & (kWave - 1);
    const int64_t s0 = xcd_span_group<G::Q>(blockIdx.x) * G::S;
    if (s0 >= N) return;  // padding workgroup: the whole block leaves
    const T* P[4] = {A, B, C, D};

    // NB chunks in flight: chunk c + NB is fetched into the buffer chunk c
    // has just left (FP64: 2; FP32: 1, a second buffer would spill)
    constexpr int NB = sizeof(T) == 8 ? 2 : 1;
    T pre[NB][4][G::EV];
#pragma unroll
    for (int c = 0; c < NB && c < G::NC; ++c)
#pragma unroll
        for (int k = 0; k < 4; ++k)
            il_fetch<T, MI>(P[k], k, c, t, n, N, stride, s0, vec != 0, pre[c][k]);

    Eq<T> e[MI];
#pragma unroll
    for (int c = 0; c < G::NC; ++c) {
        __syncthreads();  // the previous chunk has been read back
        lds_poison_wg(tile, (unsigned)sizeof(tile));
        {
            const int r = t / G::VPR;
            const int sb = (t % G::VPR) * G::EV;
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int q = 0; q < G::EV; ++q)
                    tile[k * G::AP + il_addr<T, MI>(sb + q, r)] = pre[c % NB][k][q];
        }
        __syncthreads();
        if (c + NB < G::NC) {
#pragma unroll
            for (int k = 0; k < 4; ++k)
                il_fetch<T, MI>(P[k], k, c + NB, t, n, N, stride, s0, vec != 0, pre[c % NB][k]);
        }
        if (lane / G::LC == c) {
            const int base = w * G::SP + (lane % G::LC) * G::GP;
            T v[4][MI];
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int q = 0; q < MI / G::EV; ++q) {
                    const u32x4 x4 =
                        *reinterpret_cast<const u32x4*>(&tile[k * G::AP + base + q * G::EV]);
                    __builtin_memcpy(&v[k][q * G::EV], &x4, 16);
                }
#pragma unroll
            for (int i = 0; i < MI; ++i) e[i] = Eq<T>{v[0][i], v[1][i], v[2][i], v[3][i]};
        }
    }

    T xv[MI];
    wave_core<T, MI, 1>(e, xv, lane, lane, kWave, rest);

    for (int c = 0; c < G::NC; ++c) {
        __syncthreads();
        lds_poison_wg(tile, (unsigned)sizeof(tile));
        if (lane / G::LC == c) {
            const int base = w * G::SP + (lane % G::LC) * G::GP;
#pragma unroll
            for (int q = 0; q < MI / G::EV; ++q) {
                u32x4 x4;
                __builtin_memcpy(&x4, &xv[q * G::EV], 16);
                *reinterpret_cast<u32x4*>(&tile[base + q * G::EV]) = x4;
            }
        }
        __syncthreads();
        const int r = t / G::VPR;
        const int sb = (t % G::VPR) * G::EV;
        const int64_t row = (int64_t)c * G::RC + r;
        const int64_t s = s0 + sb;
        if (row < n) {
            T v[G::EV];
#pragma unroll
            for (int q = 0; q < G::EV; ++q) v[q] = tile[il_addr<T, MI>(sb + q, r)];
            const int64_t off = row * stride + s;
            if (vec && s + G::EV <= N) {
                u32x4 x4;
                __builtin_memcpy(&x4, v, 16);
                gstore16<false>(X + off, x4);
            } else {
#pragma unroll
                for (int q = 0; q < G::EV; ++q)
                    if (s + q < N) X[off + q] = v[q];
            }
        }
    }
}

bool il_aligned(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

}  // namespace

template <typename T, int MI>
hipError_t launch_il_m(const InterleavedArgs<T>& g)
{
    using G = IlGeom<T, MI>;
    const int64_t groups = (g.N + G::S - 1) / G::S;
    constexpr int64_t kMul = 8 * G::Q;  // xcd_span_group is one-to-one on multiples of this
    const int64_t grid = (groups + kMul - 1) / kMul * kMul;
    if (grid > 0x7fffffffLL) return hipErrorInvalidConfiguration;
    const int vec = (g.stride % G::EV == 0 && il_aligned(g.a) && il_aligned(g.b) &&
                     il_aligned(g.c) && il_aligned(g.d) && il_aligned(g.x))
                        ? 1
                        : 0;
    hipLaunchKernelGGL((il_reg_kernel<T, MI>), dim3((unsigned)grid), dim3(G::S * kWave), 0,
                       g.stream, g.a, g.b, g.c, g.d, g.x, g.n, g.N, g.stride, kWave, vec);
    return hipGetLastError();
}

// One wave per system: 16, 8 or 4 unknowns per lane for n up to 1024, 512
// or 256; shorter systems stay with the lane-per-system elimination. The
// bounds are measured (2^26 unknowns, against that elimination): FP32
// n = 512: 352 / 492 us, n = 256 at 4 per lane: 536 / 469 us (also with 8
// systems per workgroup), so FP32 starts above 256; FP64
// n = 512: 647 / 945 us, n = 200: 622 / 984 us, n = 129: 1012 / 1069 us.
template <typename T>
bool interleaved_reg_takes(int64_t n, int64_t N)
{
    const int64_t lo = sizeof(T) == 4 ? kWave * 4 : kWave * 2;
    if (n <= lo || n > kWave * 16 || N <= 0) return false;
    return (N + 15) / 16 <= 0x7fffffffLL / 64;
}

template <typename T>
hipError_t launch_interleaved_reg(const InterleavedArgs<T>& g, bool* taken)
{
    *taken = false;
    if (!interleaved_reg_takes<T>(g.n, g.N)) return hipSuccess;
    *taken = true;
    if (g.n > kWave * 8) return launch_il_m<T, 16>(g);
    if (g.n > kWave * 4) return launch_il_m<T, 8>(g);
    return launch_il_m<T, 4>(g);
}

template bool interleaved_reg_takes<float>(int64_t, int64_t);
template bool interleaved_reg_takes<double>(int64_t, int64_t);
template hipError_t launch_interleaved_reg<float>(const InterleavedArgs<float>&, bool*);
template hipError_t launch_interleaved_reg<double>(const InterleavedArgs<double>&, bool*);

}  // namespace trid
