// kern_part.hip -- CR / CR-PCR for systems longer than the register kernels
// hold (n above 16,384 unknowns in FP32, 8,192 in FP64), by separators.
//
// No counterpart in the reference, whose cr / crpcr kernels keep a whole
// system in one block's shared memory (cr_app.cu:37-38) and so end at a few
// thousand unknowns; pcrGauss is its only form for long systems.
//
// A system is cut into chunks of 1,024 rows, one wave each (16 unknowns per
// lane, the layout of the register core). The last row of chunk k,
// p_k = 1024 k + 1023, is its separator. CR's first ten forward levels over
// the whole system leave exactly the separators, coupled to their
// neighbours: a tridiagonal system of K = ceil(n / 1024) unknowns per system.
// The reduced equation of p_k is a chain of ten reduce steps whose left
// operands are rows of chunk k and whose right operands are the "prefix"
// rows of chunk k + 1 (rows 1024 (k+1) + 2^j - 1 after j levels, which depend
// on that chunk alone). So:
//
//   pass 1  (part_reduce)  each wave loads its chunk, runs the ten CR levels
//           and writes 21 equations: the separator's own row, the ten left
//           operands of its chain, the chunk's ten prefix rows;
//   pass 2  (part_sep)     one thread per separator finishes its chain from
//           its record and the next chunk's; the separator system (K
//           unknowns, contiguous per system) is solved by the register
//           kernels, or by this form again when K is itself too long;
//   pass 3  (part_solve)   each wave loads its chunk again, folds the left
//           separator's value into its first row, replaces its own separator
//           row by x = value, and solves the chunk with the register core.
//
// HBM traffic per unknown: 4 T read (pass 1) + 4 T read, 1 T written
// (pass 3), plus 21 equations and 5 separator values per 1,024 unknowns:
// 9 T + 0.35 bytes, against the 5 T of a solve that fits a workgroup.
// Rows past n are identity rows (x = 0), so any n works. System s starts at
// s * bs (bs = n: a contiguous batch; bs > n: a strided one, the elements
// between systems are neither read nor written). Loads and stores are
// 16-byte vectors staged through LDS when n and bs are multiples of 16 bytes
// and the arrays are 16-byte aligned, element-wise otherwise.
#include "reg_core.hpp"
#include "reg_io.hpp"
#include "trid_device.hpp"
#include "trid_launch.hpp"
#include "trid_plan.hpp"

namespace trid {

namespace {

constexpr int kPM = 16;                // unknowns per lane
constexpr int kChunk = kWave * kPM;    // rows per chunk
constexpr int kLevels = 10;            // log2(kChunk): CR levels of pass 1
constexpr int kRec = 1 + 2 * kLevels;  // equations recorded per chunk

// Record of chunk c: 21 equations, entry j = 0: the separator row; 1..10:
// left operand of chain level j; 11 + j: prefix row after j levels
// (j = 0..9). Chunk-major: one chunk's record is one contiguous block,
// written by one store instruction of its wave.
template <typename T>
__device__ __forceinline__ void put_eq(T* p, const Eq<T>& e)
{
    constexpr int V = (int)sizeof(Eq<T>) / 16;
    u32x4 v[V];
    __builtin_memcpy(v, &e, sizeof(Eq<T>));
#pragma unroll
    for (int i = 0; i < V; ++i) reinterpret_cast<u32x4*>(p)[i] = v[i];
}

template <typename T>
__device__ __forceinline__ Eq<T> get_eq(const T* p)
{
    constexpr int V = (int)sizeof(Eq<T>) / 16;
    u32x4 v[V];
#pragma unroll
    for (int i = 0; i < V; ++i) v[i] = reinterpret_cast<const u32x4*>(p)[i];
    Eq<T> e;
    __builtin_memcpy(&e, v, sizeof(Eq<T>));
    return e;
}

// Rows of chunk `chunk` into lane-major registers (lane l: rows 16 l ..
// 16 l + 15 of the chunk), identity past the system's end; a of the
// system's first row and c of its last row are ignored, as everywhere.
template <typename T>
__device__ __forceinline__ void part_load(const T* __restrict__ A, const T* __restrict__ B,
                                          const T* __restrict__ C, const T* __restrict__ D,
                                          int64_t n, int64_t bs, int64_t K, int64_t chunk,
                                          int lane, char* lds, bool vec, Eq<T> (&e)[kPM],
                                          int64_t& base, int64_t& rows, int64_t& k)
{
    const int64_t sys = chunk / K;
    k = chunk - sys * K;
    const int64_t r0 = k * kChunk;
    base = sys * bs + r0;
    rows = n - r0 < kChunk ? n - r0 : kChunk;
    if (vec) {
        load_staged<T, kPM, true>(A, B, C, D, base, base + rows, lane, lds, e,
                                  ConstCoef<T>{false, T(0), T(0), T(0)});
    } else {
#pragma unroll
        for (int s = 0; s < kPM; ++s) {
            const int64_t j = (int64_t)lane * kPM + s;
            if (j < rows)
                e[s] = Eq<T>{A[base + j], B[base + j], C[base + j], D[base + j]};
        }
    }
#pragma unroll
    for (int s = 0; s < kPM; ++s) {
        const int64_t j = (int64_t)lane * kPM + s;
        if (j >= rows) e[s] = identity_eq<T>();
        if (r0 + j == n - 1) e[s].c = T(0);
    }
    if (k == 0 && lane == 0) e[0].a = T(0);
}

template <typename T>
__global__ __launch_bounds__(256) void part_reduce_kernel(const T* __restrict__ A,
                                                          const T* __restrict__ B,
                                                          const T* __restrict__ C,
                                                          const T* __restrict__ D, int64_t n,
                                                          int64_t bs, int64_t K, int64_t chunks,
                                                          int vec, T* __restrict__ rec)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t chunk = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    if (chunk >= chunks) return;  // whole waves
    extern __shared__ __align__(16) char trid_smem[];
    char* lds = trid_smem + (threadIdx.x / kWave) * Stage<T, kPM>::kWaveBytes;
    Eq<T> e[kPM];
    int64_t base, rows, k;
    part_load<T>(A, B, C, D, n, bs, K, chunk, lane, lds, vec != 0, e, base, rows, k);

    // the record is gathered in the wave's staging area (free once the
    // chunk is loaded), then leaves as one contiguous block
    T* rl = reinterpret_cast<T*>(lds);
    __builtin_amdgcn_wave_barrier();
    lds_poison_wave(lds, Stage<T, kPM>::kWaveBytes, lane);
    if (lane == kWave - 1) put_eq(rl, e[kPM - 1]);
    if (lane == 0) put_eq(rl + 4 * 11, e[0]);

    // in-lane CR levels 1..4 (the top slot of lane 63, the separator, has
    // no right neighbour here: its chain is finished in pass 2)
#pragma unroll
    for (int lv = 1; lv <= Log2<kPM>::value; ++lv) {
        const int SS = 1 << lv, H = SS >> 1;
        if (lane == kWave - 1) put_eq(rl + 4 * lv, e[kPM - 1 - H]);
        const Eq<T> nb = lane_step_eq<1>(e[H - 1], lane + 1 < kWave);
#pragma unroll
        for (int j = 0; j < kPM / SS; ++j) {
            const int i = SS * j + SS - 1;
            const int ir = (i + H < kPM) ? i + H : i;
            const Eq<T> rgt = (i + H < kPM) ? e[ir] : nb;
            e[i] = reduce_eq(e[i], e[i - H], rgt);
        }
        if (lane == 0) put_eq(rl + 4 * (11 + lv), e[SS - 1]);
    }
    // cross-lane CR levels 5..10 on the top slots
#pragma unroll
    for (int q = 1; q <= 6; ++q) {
        const int hs = 1 << (q - 1);
        const Eq<T> lft = lane_read_eq(e[kPM - 1], lane - hs, lane - hs >= 0);
        const Eq<T> rgt = lane_read_eq(e[kPM - 1], lane + hs, lane + hs < kWave);
        if (lane == kWave - 1) put_eq(rl + 4 * (4 + q), lft);
        if (q == 6) break;  // the separator's own level: pass 2
        const Eq<T> r = reduce_eq(e[kPM - 1], lft, rgt);
        if (((lane + 1) & (2 * hs - 1)) == 0) e[kPM - 1] = r;
        if (lane == (1 << q) - 1) put_eq(rl + 4 * (15 + q), e[kPM - 1]);
    }
    __builtin_amdgcn_wave_barrier();
    constexpr int kPieces = kRec * (int)sizeof(Eq<T>) / 16;  // 16-byte pieces
    if (lane < kPieces)
        reinterpret_cast<u32x4*>(rec + chunk * kRec * 4)[lane] =
            reinterpret_cast<const u32x4*>(rl)[lane];
}

// One thread per separator: the chain of its row through the ten levels,
// right operands from the next chunk of the same system (none for the last).
template <typename T>
__global__ __launch_bounds__(256) void part_sep_kernel(const T* __restrict__ rec, int64_t K,
                                                       int64_t chunks, T* __restrict__ sa,
                                                       T* __restrict__ sb, T* __restrict__ sc,
                                                       T* __restrict__ sd)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= chunks) return;
    const T* me = rec + c * kRec * 4;
    const T* nx = me + kRec * 4;
    const bool has_next = (c % K) + 1 < K;
    Eq<T> top = get_eq(me);
#pragma unroll
    for (int lv = 1; lv <= kLevels; ++lv) {
        const Eq<T> lft = get_eq(me + 4 * lv);
        const Eq<T> rgt = has_next ? get_eq(nx + 4 * (11 + lv - 1)) : identity_eq<T>();
        top = reduce_eq(top, lft, rgt);
    }
    sa[c] = top.a;
    sb[c] = top.b;
    sc[c] = top.c;
    sd[c] = top.d;
}

template <typename T, int R>
__global__ __launch_bounds__(256) void part_solve_kernel(const T* __restrict__ A,
                                                         const T* __restrict__ B,
                                                         const T* __restrict__ C,
                                                         const T* __restrict__ D,
                                                         T* __restrict__ X, int64_t n, int64_t bs,
                                                         int64_t K, int64_t chunks, int vec,
                                                         int rest, const T* __restrict__ xs)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t chunk = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    if (chunk >= chunks) return;
    extern __shared__ __align__(16) char trid_smem[];
    char* lds = trid_smem + (threadIdx.x / kWave) * Stage<T, kPM>::kWaveBytes;
    // separator values first, so that their loads overlap the chunk's
    const bool first_of_sys = chunk % K == 0;
    const T xl = (lane == 0 && !first_of_sys) ? xs[chunk - 1] : T(0);
    const T xr = lane == kWave - 1 ? xs[chunk] : T(0);
    Eq<T> e[kPM];
    int64_t base, rows, k;
    part_load<T>(A, B, C, D, n, bs, K, chunk, lane, lds, vec != 0, e, base, rows, k);
    // the left separator's value enters the first row's right-hand side;
    // the chunk's own separator row becomes x = its value
    if (lane == 0) {
        e[0].d -= e[0].a * xl;
        e[0].a = T(0);
    }
    if (lane == kWave - 1) e[kPM - 1] = Eq<T>{T(0), T(1), T(0), xr};

    T xv[kPM];
    wave_core<T, kPM, R, true>(e, xv, lane, lane, kWave, rest);

    if (vec) {
        store_staged<T, kPM, true>(X, base, base + rows, lane, lds, xv);
    } else {
#pragma unroll
        for (int s = 0; s < kPM; ++s) {
            const int64_t j = (int64_t)lane * kPM + s;
            if (j < rows) X[base + j] = xv[s];
        }
    }
}

size_t round256(size_t v) { return (v + 255) / 256 * 256; }

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// Plan of the separator system: N systems of K unknowns, same method,
// default switch size.
int sep_plan(int method, int64_t K, int64_t N, int elem, Plan* p)
{
    trid_opts o{};
    o.rest = method == TRID_CR ? 2 : 0;
    return make_plan(method, K, N, elem, o, p);
}

}  // namespace

int64_t part_chunks(int64_t n) { return (n + kChunk - 1) / kChunk; }

template <typename T>
size_t part_ws_bytes(int method, int64_t n, int64_t N)
{
    const int64_t K = part_chunks(n);
    const size_t chunks = (size_t)K * (size_t)N;
    size_t bytes = round256(chunks * kRec * 4 * sizeof(T)) + 5 * round256(chunks * sizeof(T));
    Plan p;
    if (sep_plan(method, K, N, (int)sizeof(T), &p) == TRID_OK && p.family == Family::Part)
        bytes += part_ws_bytes<T>(method, K, N);
    return bytes;
}

template <typename T>
hipError_t launch_part(const PartArgs<T>& g)
{
    const int64_t K = part_chunks(g.n);
    const int64_t chunks = K * g.N;
    if (chunks <= 0) return hipSuccess;
    const int64_t blocks = (chunks * kWave + 255) / 256;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidConfiguration;
    const int R = g.rest > kWave ? g.rest / kWave : 1;
    if (g.rest < 2 || (g.rest > kWave && g.rest % kWave != 0) || R > 8)
        return hipErrorInvalidValue;

    char* ws = static_cast<char*>(g.ws);
    T* rec = reinterpret_cast<T*>(ws);
    ws += round256((size_t)chunks * kRec * 4 * sizeof(T));
    T* sep[5];
    for (int i = 0; i < 5; ++i) {
        sep[i] = reinterpret_cast<T*>(ws);
        ws += round256((size_t)chunks * sizeof(T));
    }
    constexpr int E = 16 / (int)sizeof(T);
    const int64_t bs = g.bs ? g.bs : g.n;
    const int vec = (g.n % E == 0 && bs % E == 0 && aligned16(g.a) && aligned16(g.b) &&
                     aligned16(g.c) && aligned16(g.d) && aligned16(g.x))
                        ? 1
                        : 0;
    // pass 1 gathers its record in the staging area on either load path
    const size_t lds1 = 4 * (size_t)Stage<T, kPM>::kWaveBytes;
    const size_t lds = vec ? lds1 : 0;

    hipLaunchKernelGGL((part_reduce_kernel<T>), dim3((unsigned)blocks), dim3(256), lds1, g.stream,
                       g.a, g.b, g.c, g.d, g.n, bs, K, chunks, vec, rec);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((part_sep_kernel<T>), dim3((unsigned)((chunks + 255) / 256)), dim3(256),
                       0, g.stream, rec, K, chunks, sep[0], sep[1], sep[2], sep[3]);
    if ((e = hipGetLastError()) != hipSuccess) return e;

    // the separator system
    Plan p;
    if (sep_plan(g.method, K, g.N, (int)sizeof(T), &p) != TRID_OK) return hipErrorInvalidValue;
    if (p.family == Family::Reg) {
        RegArgs<T> r{};
        r.a = sep[0];
        r.b = sep[1];
        r.c = sep[2];
        r.d = sep[3];
        r.x = sep[4];
        r.n = K;
        r.N = g.N;
        r.M = p.M;
        r.L = p.L;
        r.rest = p.rest;
        r.io = p.io;
        r.vec = p.dense;  // workspace arrays are 256-byte aligned
        r.stream = g.stream;
        e = r.L <= kWave ? launch_reg_crpcr<T>(r) : launch_reg_mw<T>(r);
    } else if (p.family == Family::Part) {
        PartArgs<T> q = g;
        q.a = sep[0];
        q.b = sep[1];
        q.c = sep[2];
        q.d = sep[3];
        q.x = sep[4];
        q.n = K;
        q.bs = 0;
        q.rest = p.rest;
        q.ws = ws;
        e = launch_part<T>(q);
    } else {
        return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;

#define TRID_PART_CASE(RR)                                                                     \
    case RR:                                                                                   \
        hipLaunchKernelGGL((part_solve_kernel<T, RR>), dim3((unsigned)blocks), dim3(256), lds, \
                           g.stream, g.a, g.b, g.c, g.d, g.x, g.n, bs, K, chunks, vec, g.rest,  \
                           sep[4]);                                                            \
        break;
    switch (R) {
        TRID_PART_CASE(1)
        TRID_PART_CASE(2)
        TRID_PART_CASE(4)
        TRID_PART_CASE(8)
        default: return hipErrorInvalidValue;
    }
#undef TRID_PART_CASE
    return hipGetLastError();
}

template size_t part_ws_bytes<float>(int, int64_t, int64_t);
template size_t part_ws_bytes<double>(int, int64_t, int64_t);
template hipError_t launch_part<float>(const PartArgs<float>&);
template hipError_t launch_part<double>(const PartArgs<double>&);

}  // namespace trid
