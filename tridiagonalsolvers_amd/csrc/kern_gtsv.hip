// kern_gtsv.hip -- batched tridiagonal solve with partial pivoting (row
// interchanges: the LAPACK ?gtsv algorithm for one right-hand side), for
// systems that are not diagonally dominant.
//
// Reference counterpart: tridiagonalMKL.cpp:5-26 calls MKL's ?gtsv on each
// system in turn, passing &a[s*n+1] as the sub-diagonal so that a[first] is
// skipped (tridiagonalMKL.cpp:16-18).
//
// Elimination of the fresh row i+1 (l = a_{i+1}, b_{i+1}, c_{i+1}, d_{i+1}) by
// the current row i, whose values after the earlier steps are (D, U, B):
//   |D| >= |l|  no interchange: f = l / D; b_{i+1} -= f U; d_{i+1} -= f B.
//               Factored row i = (D, U, 0, B). D == 0 (then l == 0 as well)
//               is a zero pivot: info = i+1 and nothing is updated.
//   otherwise   interchange: f = D / l. Factored row i = (l, b_{i+1},
//               c_{i+1}, d_{i+1}) -- its third entry W is the fill-in on the
//               second super-diagonal -- and row i+1 becomes
//               (U - f b_{i+1}, -f c_{i+1}, B - f d_{i+1}).
// D == 0 on the last row gives info = n. Back substitution:
//   x_{n-1} = B/D,  x_{n-2} = (B - U x_{n-1})/D,
//   x_i = ((B - U x_{i+1}) - W x_{i+2}) / D.
// These are LAPACK's operations in LAPACK's order. Contraction is off and
// the divisions are IEEE, so every system's x equals that of a scalar CPU
// restatement bit for bit. A system with info != 0 gets x = NaN (LAPACK
// leaves its solution uncomputed).
//
// Parallel form: one lane per system, 64 systems per wave, one wave per
// workgroup. Each system is one serial chain, so the kernel is built to
// stream rather than to share work:
// * pass 1 reads every row once and keeps only a checkpoint of the running
//   row (D, U, B) at the start of each chunk of rows;
// * pass 2 walks the chunks backwards: it reads a chunk again, re-runs the
//   elimination from the chunk's checkpoint (the same operations, so the
//   same bits as pass 1) with the factored rows held in registers,
//   back-substitutes them and writes x.
// HBM traffic is 9 elements per unknown (4 + 4 read, 1 written) plus a
// 4-element checkpoint per chunk and system, where keeping the factored rows
// in memory would cost 13 (4 read, 4 written and read back, 1 written).
// * System-major data (element i of system s at s*n + i): a chunk is 128
//   bytes of each system; eight lanes read one system's chunk as 16-byte
//   vectors (8 systems per load), the chunk is transposed in the wave's LDS
//   (row pitch 66 elements: the transposing writes and the per-lane reads
//   hit distinct banks) and x goes back out the same way.
// * Interleaved data (element i of system s at i*stride + s): each lane reads
//   and writes its own system directly; every access is coalesced across the
//   wave's 64 systems. Chunks of 32 rows (FP32) or 16 (FP64).
#include "reg_io.hpp"
#include "trid_device.hpp"
#include "trid_launch.hpp"

#pragma clang fp contract(off)

// Bytes of one system per array in a system-major chunk (64 or 128).
// (128: a whole cache line per system and array; on 65,536 x 1,024 it runs
// 16% (FP32) and 21% (FP64) faster than 64, more bytes being in flight per
// wave at one wave per SIMD.)
#ifndef TRID_GTSV_CHUNK_BYTES
#define TRID_GTSV_CHUNK_BYTES 128
#endif
// Rows per interleaved chunk, per element type: the most that stay in
// registers (chunk in flight + chunk in use + factored rows; one wave per
// SIMD leaves each wave the whole register file). On 65,536 x 1,024: FP32
// 16 rows 0.689 ms, 32 rows 0.589 ms (40 spills); FP64 16 rows 1.007 ms,
// 32 rows 2.51 ms (24 already spills).
#ifndef TRID_GTSV_IL_ROWS_F32
#define TRID_GTSV_IL_ROWS_F32 32
#endif
#ifndef TRID_GTSV_IL_ROWS_F64
#define TRID_GTSV_IL_ROWS_F64 16
#endif

namespace trid {

namespace {

template <typename T>
struct PivRow {  // a fresh row: a_i, b_i, c_i, d_i
    T l, b, c, d;
};

template <typename T>
struct PivFac {  // a factored row: diagonal, first and second super-diagonal, rhs
    T D, U, W, B;
};

__device__ __forceinline__ float pabs(float v) { return __builtin_fabsf(v); }
__device__ __forceinline__ double pabs(double v) { return __builtin_fabs(v); }
template <typename T>
__device__ __forceinline__ T pnan();
template <>
__device__ __forceinline__ float pnan<float>() { return __builtin_nanf(""); }
template <>
__device__ __forceinline__ double pnan<double>() { return __builtin_nan(""); }

// One elimination step: (dg, up, rh) is the current row i, f the fresh row
// i+1. Writes the factored row i and leaves row i+1 in (dg, up, rh).
// Returns true on a zero pivot. Both cases are one division and two
// multiply-subtracts on selected operands (lanes of a wave take different
// cases; selects keep them in step):
//   no interchange: m = l/D,  b' = b - m U,  d' = d - m B,  c' = c
//   interchange:    m = D/l,  b' = U - m b,  d' = B - m d,  c' = -(m c)
template <typename T>
__device__ __forceinline__ bool pivot_step(T& dg, T& up, T& rh, const PivRow<T>& f,
                                           PivFac<T>& out)
{
    const bool swap = !(pabs(dg) >= pabs(f.l));
    const bool zero = !swap && dg == T(0);
    const T m = (swap ? dg : f.l) / (swap ? f.l : dg);
    const T nb = (swap ? up : f.b) - m * (swap ? f.b : up);
    const T nd = (swap ? rh : f.d) - m * (swap ? f.d : rh);
    const T nc = swap ? -(m * f.c) : f.c;
    out = swap ? PivFac<T>{f.l, f.b, f.c, f.d} : PivFac<T>{dg, up, T(0), rh};
    dg = zero ? f.b : nb;
    up = nc;
    rh = zero ? f.d : nd;
    return zero;
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- system-major chunks, transposed through LDS ------------------------

template <typename T, bool VEC>
struct SysMajorIO {
    static constexpr int kVec = 16 / (int)sizeof(T);                      // elements per vector
    static constexpr int kRows = TRID_GTSV_CHUNK_BYTES / (int)sizeof(T);  // rows per chunk
    static constexpr int kLanesPerSys = kRows / kVec;                     // 4 or 8
    static constexpr int kSysPerLoad = kWave / kLanesPerSys;              // 16 or 8
    static constexpr int kLoads = kWave / kSysPerLoad;                    // 4 or 8
    // LDS row pitch: the lanes of one transposing write hold rows
    // rl = j*kVec (j < kLanesPerSys) of kSysPerLoad consecutive systems; a
    // pitch of 64 + 16/kLanesPerSys elements puts the kLanesPerSys groups
    // on disjoint banks.
    static constexpr int kPitch = kWave + 16 / kLanesPerSys;
    static constexpr int kLdsElems = 4 * kRows * kPitch;
    static_assert(kLanesPerSys == 4 || kLanesPerSys == 8, "64- or 128-byte chunks");

    const T* p[4];
    T* X;
    int64_t n, N, sys0;
    int lane;
    T* tile;  // kLdsElems elements of this wave
    u32x4 pre[4][kLoads];

    __device__ __forceinline__ int sl(int q) const { return q * kSysPerLoad + lane / kLanesPerSys; }
    __device__ __forceinline__ int rl() const { return (lane % kLanesPerSys) * kVec; }

    __device__ __forceinline__ void prefetch(int64_t c)
    {
        const int64_t row0 = c * kRows;
        const int rows = (int)(n - row0 < kRows ? n - row0 : kRows);
#pragma unroll
        for (int q = 0; q < kLoads; ++q) {
            const int64_t sys = sys0 + sl(q);
            const bool live = sys < N;
            const int64_t off = sys * n + row0 + rl();
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if constexpr (VEC) {
                    pre[k][q] = (live && rl() < rows) ? gload16<false>(p[k] + off)
                                                      : u32x4{0u, 0u, 0u, 0u};
                } else {
                    T t[kVec];
#pragma unroll
                    for (int e = 0; e < kVec; ++e)
                        t[e] = (live && rl() + e < rows) ? p[k][off + e] : T(0);
                    __builtin_memcpy(&pre[k][q], t, 16);
                }
            }
        }
    }

    // The prefetched chunk becomes the current one (readable with row()).
    __device__ __forceinline__ void land()
    {
        wave_sync();
#pragma unroll
        for (int q = 0; q < kLoads; ++q) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                T t[kVec];
                __builtin_memcpy(t, &pre[k][q], 16);
#pragma unroll
                for (int e = 0; e < kVec; ++e) tile[(k * kRows + rl() + e) * kPitch + sl(q)] = t[e];
            }
        }
        wave_sync();
    }

    __device__ __forceinline__ PivRow<T> row(int r) const
    {
        return PivRow<T>{tile[(0 * kRows + r) * kPitch + lane], tile[(1 * kRows + r) * kPitch + lane],
                         tile[(2 * kRows + r) * kPitch + lane], tile[(3 * kRows + r) * kPitch + lane]};
    }

    __device__ __forceinline__ void store_x(int64_t c, const T (&xs)[kRows])
    {
        const int64_t row0 = c * kRows;
        const int rows = (int)(n - row0 < kRows ? n - row0 : kRows);
        wave_sync();
#pragma unroll
        for (int r = 0; r < kRows; ++r) tile[r * kPitch + lane] = xs[r];
        wave_sync();
#pragma unroll
        for (int q = 0; q < kLoads; ++q) {
            const int64_t sys = sys0 + sl(q);
            const bool live = sys < N;
            const int64_t off = sys * n + row0 + rl();
            T t[kVec];
#pragma unroll
            for (int e = 0; e < kVec; ++e) t[e] = tile[(rl() + e) * kPitch + sl(q)];
            if constexpr (VEC) {
                if (live && rl() < rows) {
                    u32x4 v;
                    __builtin_memcpy(&v, t, 16);
                    gstore16<false>(X + off, v);
                }
            } else {
#pragma unroll
                for (int e = 0; e < kVec; ++e)
                    if (live && rl() + e < rows) X[off + e] = t[e];
            }
        }
    }
};

// ---- interleaved chunks, one lane per system, no LDS --------------------

template <typename T>
struct InterleavedIO {
    static constexpr int kRows = sizeof(T) == 4 ? TRID_GTSV_IL_ROWS_F32 : TRID_GTSV_IL_ROWS_F64;
    static constexpr int kLdsElems = 0;

    const T* p[4];
    T* X;
    int64_t n, N, sys0, stride;
    int lane;
    T* tile;  // unused
    T pre[4][kRows];
    T cur[4][kRows];

    __device__ __forceinline__ void prefetch(int64_t c)
    {
        const int64_t row0 = c * kRows;
        const int64_t sys = sys0 + lane;
        const bool live = sys < N;
#pragma unroll
        for (int r = 0; r < kRows; ++r) {
            const bool in = live && row0 + r < n;
            const int64_t off = (row0 + r) * stride + sys;
#pragma unroll
            for (int k = 0; k < 4; ++k) pre[k][r] = in ? p[k][off] : T(0);
        }
    }

    __device__ __forceinline__ void land()
    {
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int r = 0; r < kRows; ++r) cur[k][r] = pre[k][r];
    }

    __device__ __forceinline__ PivRow<T> row(int r) const
    {
        return PivRow<T>{cur[0][r], cur[1][r], cur[2][r], cur[3][r]};
    }

    __device__ __forceinline__ void store_x(int64_t c, const T (&xs)[kRows])
    {
        const int64_t row0 = c * kRows;
        const int64_t sys = sys0 + lane;
        if (sys >= N) return;
#pragma unroll
        for (int r = 0; r < kRows; ++r)
            if (row0 + r < n) X[(row0 + r) * stride + sys] = xs[r];
    }
};

// Checkpoint records: 4 elements (D, U, B, 0) per lane and chunk, 64 lanes
// contiguous, so each record store or load covers 1 or 2 KiB.
template <typename T>
__device__ __forceinline__ void ck_store(T* ck, int lane, T dg, T up, T rh)
{
    T t[4] = {dg, up, rh, T(0)};
    u32x4 v[sizeof(T) / 4];
    __builtin_memcpy(v, t, sizeof(t));
#pragma unroll
    for (int h = 0; h < (int)(sizeof(T) / 4); ++h)
        gstore16<false>(ck + (int64_t)lane * 4 + h * (16 / (int)sizeof(T)), v[h]);
}

template <typename T>
struct CkRec {
    u32x4 v[sizeof(T) / 4];
};

template <typename T>
__device__ __forceinline__ CkRec<T> ck_load(const T* ck, int lane)
{
    CkRec<T> r;
#pragma unroll
    for (int h = 0; h < (int)(sizeof(T) / 4); ++h)
        r.v[h] = gload16<false>(ck + (int64_t)lane * 4 + h * (16 / (int)sizeof(T)));
    return r;
}

template <typename T, class IO>
__device__ __forceinline__ void gtsv_wave(IO& io, int64_t n, T* __restrict__ ck,
                                          int32_t* __restrict__ info)
{
    constexpr int R = IO::kRows;
    const int lane = io.lane;
    const int64_t nch = (n + R - 1) / R;
    const int64_t ck_chunk = (int64_t)kWave * 4;  // elements per chunk record block

    // pass 1: forward elimination, checkpoints only
    T dg = T(0), up = T(0), rh = T(0);
    int64_t bad = 0;
    io.prefetch(0);
    for (int64_t c = 0; c < nch; ++c) {
        io.land();
        if (c + 1 < nch) io.prefetch(c + 1);
        const int64_t row0 = c * R;
        const int rows = (int)(n - row0 < R ? n - row0 : R);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r < rows) {
                const int64_t i = row0 + r;
                PivRow<T> f = io.row(r);
                if (i == n - 1) f.c = T(0);
                if (i == 0) {
                    dg = f.b;
                    up = f.c;
                    rh = f.d;
                } else {
                    PivFac<T> o;
                    if (pivot_step(dg, up, rh, f, o) && bad == 0) bad = i;
                }
                if (r == 0) ck_store(ck + c * ck_chunk, lane, dg, up, rh);
            }
        }
    }
    if (dg == T(0) && bad == 0) bad = n;
    const int64_t sys = io.sys0 + lane;
    if (info && sys < io.N) info[sys] = (int32_t)(bad > INT32_MAX ? INT32_MAX : bad);

    // pass 2: chunks in reverse, re-eliminate from the checkpoint, substitute
    T x1 = T(0), x2 = T(0);
    PivRow<T> nxt{T(0), T(0), T(0), T(0)};
    io.prefetch(nch - 1);
    CkRec<T> ckp = ck_load(ck + (nch - 1) * ck_chunk, lane);
    for (int64_t c = nch - 1; c >= 0; --c) {
        io.land();
        T t[4];
        __builtin_memcpy(t, ckp.v, sizeof(t));
        dg = t[0];
        up = t[1];
        rh = t[2];
        if (c > 0) {
            io.prefetch(c - 1);
            ckp = ck_load(ck + (c - 1) * ck_chunk, lane);
        }
        const int64_t row0 = c * R;
        const int rows = (int)(n - row0 < R ? n - row0 : R);
        PivFac<T> F[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r < rows) {
                const int64_t i = row0 + r;
                if (i == n - 1) {
                    F[r] = PivFac<T>{dg, up, T(0), rh};
                } else {
                    PivRow<T> f = (r + 1 < rows) ? io.row(r + 1 < R ? r + 1 : R - 1) : nxt;
                    if (i + 1 == n - 1) f.c = T(0);
                    (void)pivot_step(dg, up, rh, f, F[r]);
                }
            }
        }
        T xs[R];
#pragma unroll
        for (int r = R - 1; r >= 0; --r) {
            xs[r] = T(0);
            if (r < rows) {
                const int64_t i = row0 + r;
                T x;
                if (i == n - 1) x = F[r].B / F[r].D;
                else if (i == n - 2) x = (F[r].B - F[r].U * x1) / F[r].D;
                else x = ((F[r].B - F[r].U * x1) - F[r].W * x2) / F[r].D;
                x2 = x1;
                x1 = x;
                xs[r] = bad ? pnan<T>() : x;
            }
        }
        nxt = io.row(0);
        io.store_x(c, xs);
    }
}

template <typename T, bool VEC>
__global__ __launch_bounds__(kWave) void gtsv_sysmajor_kernel(
    const T* __restrict__ A, const T* __restrict__ B, const T* __restrict__ C,
    const T* __restrict__ D, T* X, int64_t n, int64_t N, T* __restrict__ ck,
    int32_t* __restrict__ info)
{
    using IO = SysMajorIO<T, VEC>;
    __shared__ T tile[IO::kLdsElems];
    IO io;
    io.p[0] = A;
    io.p[1] = B;
    io.p[2] = C;
    io.p[3] = D;
    io.X = X;
    io.n = n;
    io.N = N;
    io.sys0 = (int64_t)blockIdx.x * kWave;
    io.lane = threadIdx.x;
    io.tile = tile;
    const int64_t nch = (n + IO::kRows - 1) / IO::kRows;
    gtsv_wave<T>(io, n, ck + (int64_t)blockIdx.x * nch * kWave * 4, info);
}

template <typename T>
__global__ __launch_bounds__(kWave) void gtsv_interleaved_kernel(
    const T* __restrict__ A, const T* __restrict__ B, const T* __restrict__ C,
    const T* __restrict__ D, T* X, int64_t n, int64_t N, int64_t stride, T* __restrict__ ck,
    int32_t* __restrict__ info)
{
    using IO = InterleavedIO<T>;
    IO io;
    io.p[0] = A;
    io.p[1] = B;
    io.p[2] = C;
    io.p[3] = D;
    io.X = X;
    io.n = n;
    io.N = N;
    io.stride = stride;
    io.sys0 = (int64_t)blockIdx.x * kWave;
    io.lane = threadIdx.x;
    io.tile = nullptr;
    const int64_t nch = (n + IO::kRows - 1) / IO::kRows;
    gtsv_wave<T>(io, n, ck + (int64_t)blockIdx.x * nch * kWave * 4, info);
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

}  // namespace

template <typename T>
size_t pivot_ws_bytes(int64_t n, int64_t N, bool interleaved)
{
    const int64_t rows = interleaved ? InterleavedIO<T>::kRows : SysMajorIO<T, true>::kRows;
    const int64_t nch = (n + rows - 1) / rows;
    const int64_t waves = (N + kWave - 1) / kWave;
    return (size_t)waves * (size_t)nch * kWave * 4 * sizeof(T);
}

template <typename T>
hipError_t launch_pivot(const PivotArgs<T>& g)
{
    if (g.N <= 0) return hipSuccess;
    const int64_t waves = (g.N + kWave - 1) / kWave;
    if (waves > 0x7fffffff) return hipErrorInvalidConfiguration;
    if (g.stride > 0) {
        hipLaunchKernelGGL((gtsv_interleaved_kernel<T>), dim3((unsigned)waves), dim3(kWave), 0,
                           g.stream, g.a, g.b, g.c, g.d, g.x, g.n, g.N, g.stride, g.ws, g.info);
        return hipGetLastError();
    }
    constexpr int kVec = 16 / (int)sizeof(T);
    const bool vec = (g.n % kVec) == 0 && aligned16(g.a) && aligned16(g.b) && aligned16(g.c) &&
                     aligned16(g.d) && aligned16(g.x);
    if (vec)
        hipLaunchKernelGGL((gtsv_sysmajor_kernel<T, true>), dim3((unsigned)waves), dim3(kWave), 0,
                           g.stream, g.a, g.b, g.c, g.d, g.x, g.n, g.N, g.ws, g.info);
    else
        hipLaunchKernelGGL((gtsv_sysmajor_kernel<T, false>), dim3((unsigned)waves), dim3(kWave),
                           0, g.stream, g.a, g.b, g.c, g.d, g.x, g.n, g.N, g.ws, g.info);
    return hipGetLastError();
}

template size_t pivot_ws_bytes<float>(int64_t, int64_t, bool);
template size_t pivot_ws_bytes<double>(int64_t, int64_t, bool);
template hipError_t launch_pivot<float>(const PivotArgs<float>&);
template hipError_t launch_pivot<double>(const PivotArgs<double>&);

}  // namespace trid
