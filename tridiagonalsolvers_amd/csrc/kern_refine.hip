// kern_refine.hip -- FP32 solve with FP64 residual refinement, fused per system.
//
// No counterpart in the reference (its element type is fixed at build time,
// common.h:5-12). Inputs and the solution are FP64; the solves run in FP32 on
// the register-resident CR-PCR core (reg_core.hpp); FP64 arithmetic is used
// only for the residual r = d - T x and for x += e:
//
//   x = solve32(T, d)
//   repeat `iters` times:  r = d - T x (FP64);  e = solve32(T, r);  x += e
//
// Everything stays in the lane's registers between the steps: the FP32
// operator is rebuilt from the FP64 coefficients while the residual is
// formed (the coefficients are re-read from memory two rows at a time, so
// only the FP32 equations and the FP64 unknowns stay live), and the
// neighbour unknowns the residual needs come from the adjacent lanes
// (ds_bpermute) or, in the multi-wave form, from LDS. The first pass is the
// same code with x = 0 (r = d).
#include "reg_core.hpp"
#include "trid_device.hpp"
#include "trid_launch.hpp"

namespace trid {

namespace {

// FP32 operator of the lane's rows with right-hand side r = d - T x formed
// in FP64 (x_left / x_right: the unknowns just outside the lane's block, 0
// outside the system; with x = 0 this is just d). Rows are read two at a
// time (16-byte loads on dense, aligned batches), so only the lane's FP32
// equations and FP64 unknowns stay live. Identity rows past n; a[first] and
// c[last] are ignored.
template <int M>
__device__ __forceinline__ void build_rows(const double* __restrict__ A,
                                           const double* __restrict__ B,
                                           const double* __restrict__ C,
                                           const double* __restrict__ D, const double (&x)[M],
                                           double x_left, double x_right, int64_t base,
                                           int64_t j0, int64_t n, bool live, bool vec,
                                           Eq<float> (&e)[M])
{
    constexpr int P = M >= 2 ? 2 : 1;  // rows per load
#pragma unroll
    for (int k = 0; k < M; k += P) {
        double va[P], vb[P], vc[P], vd[P];
        if (P == 2 && vec && live) {
            const int64_t o = base + j0 + k;
            const double2 ta = *reinterpret_cast<const double2*>(A + o);
            const double2 tb = *reinterpret_cast<const double2*>(B + o);
            const double2 tc = *reinterpret_cast<const double2*>(C + o);
            const double2 td = *reinterpret_cast<const double2*>(D + o);
            va[0] = ta.x; vb[0] = tb.x; vc[0] = tc.x; vd[0] = td.x;
            va[P - 1] = ta.y; vb[P - 1] = tb.y; vc[P - 1] = tc.y; vd[P - 1] = td.y;
        } else {
#pragma unroll
            for (int u = 0; u < P; ++u) {
                const int64_t j = j0 + k + u;
                const bool ok = live && j < n;
                va[u] = ok ? A[base + j] : 0.0;
                vb[u] = ok ? B[base + j] : 1.0;
                vc[u] = ok ? C[base + j] : 0.0;
                vd[u] = ok ? D[base + j] : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < P; ++u) {
            const int s = k + u;
            const int64_t j = j0 + s;
            const bool row = live && j < n;
            const double xl = s > 0 ? x[s > 0 ? s - 1 : 0] : x_left;
            const double xr = s < M - 1 ? x[s < M - 1 ? s + 1 : 0] : x_right;
            const double al = j > 0 ? va[u] : 0.0;
            const double cr = j < n - 1 ? vc[u] : 0.0;
            double r = fma(-vb[u], x[s], vd[u]);
            r = fma(-al, xl, r);
            r = fma(-cr, xr, r);
            e[s].a = row ? (float)al : 0.0f;
            e[s].b = row ? (float)vb[u] : 1.0f;
            e[s].c = row ? (float)cr : 0.0f;
            e[s].d = row ? (float)r : 0.0f;
        }
    }
}

}  // namespace

// One system per segment of L lanes of a wave (n <= 64 * M).
template <int M, int R>
__global__ __launch_bounds__(256) void refine_wave_kernel(const double* __restrict__ A,
                                                          const double* __restrict__ B,
                                                          const double* __restrict__ C,
                                                          const double* __restrict__ D,
                                                          double* __restrict__ X, int64_t n,
                                                          int64_t N, int L, int logL, int rest,
                                                          int iters, int vec)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int ll = lane & (L - 1);
    const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t sys = gtid >> logL;
    const bool live = sys < N;
    const int64_t base = sys * n;
    const int64_t j0 = (int64_t)ll * M;

    double x[M];
#pragma unroll
    for (int s = 0; s < M; ++s) x[s] = 0.0;
    Eq<float> e[M];
    float xf[M];
    for (int it = 0; it <= iters; ++it) {
        const double xl = lane_step_or_zero<-1>(x[M - 1], ll >= 1);
        const double xr = lane_step_or_zero<1>(x[0], ll + 1 < L);
        build_rows<M>(A, B, C, D, x, xl, xr, base, j0, n, live, vec != 0, e);
        wave_core<float, M, R>(e, xf, lane, ll, L, rest);
#pragma unroll
        for (int s = 0; s < M; ++s) x[s] += (double)xf[s];
    }

    if (!live) return;
#pragma unroll
    for (int s = 0; s < M; ++s)
        if (j0 + s < n) X[base + j0 + s] = x[s];
}

// One system per workgroup of W waves (n <= W * 64 * M).
template <int M>
__global__ __launch_bounds__(1024) void refine_wg_kernel(const double* __restrict__ A,
                                                         const double* __restrict__ B,
                                                         const double* __restrict__ C,
                                                         const double* __restrict__ D,
                                                         double* __restrict__ X, int64_t n,
                                                         int rest, int iters, int vec)
{
    extern __shared__ __align__(16) char trid_smem[];
    const int W = blockDim.x / kWave;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));  // uniform
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t base = (int64_t)blockIdx.x * n;
    const int64_t j0 = (int64_t)(w * kWave + lane) * M;
    // boundary unknowns of each wave, after the exchange area of wg_core
    double* edge = reinterpret_cast<double*>(trid_smem + wg_core_lds_bytes<float, M>(W));

    double x[M];
#pragma unroll
    for (int s = 0; s < M; ++s) x[s] = 0.0;
    Eq<float> e[M];
    float xf[M];
    for (int it = 0; it <= iters; ++it) {
        double xl = 0.0, xr = 0.0;
        if (it > 0) {
            // first / last unknown of every wave, for the neighbours across waves
            if (lane == 0) edge[2 * w] = x[0];
            if (lane == kWave - 1) edge[2 * w + 1] = x[M - 1];
            __syncthreads();
            const double xl0 = lane_read(x[M - 1], (lane - 1) & (kWave - 1));
            const double xr0 = lane_read(x[0], (lane + 1) & (kWave - 1));
            xl = lane > 0 ? xl0 : (w > 0 ? edge[2 * w - 1] : 0.0);
            xr = lane < kWave - 1 ? xr0 : (w + 1 < W ? edge[2 * w + 2] : 0.0);
        }
        build_rows<M>(A, B, C, D, x, xl, xr, base, j0, n, true, vec != 0, e);
        wg_core<float, M>(e, xf, w, W, lane, rest, trid_smem);
#pragma unroll
        for (int s = 0; s < M; ++s) x[s] += (double)xf[s];
    }
#pragma unroll
    for (int s = 0; s < M; ++s)
        if (j0 + s < n) X[base + j0 + s] = x[s];
}

template <int M>
static hipError_t launch_refine_wave_m(const RefineArgs& g, int R, dim3 grid, dim3 block)
{
    const int logL = __builtin_ctz((unsigned)g.L);
#define TRID_REF_CASE(RR)                                                                   \
    case RR:                                                                                \
        if constexpr (RR <= M) {                                                            \
            hipLaunchKernelGGL((refine_wave_kernel<M, RR>), grid, block, 0, g.stream, g.a,   \
                               g.b, g.c, g.d, g.x, g.n, g.N, g.L, logL, g.rest, g.iters,      \
                               g.vec);                                                      \
            return hipGetLastError();                                                       \
        }                                                                                   \
        break;
    switch (R) {
        TRID_REF_CASE(1)
        TRID_REF_CASE(2)
        TRID_REF_CASE(4)
        TRID_REF_CASE(8)
        TRID_REF_CASE(16)
        default: break;
    }
#undef TRID_REF_CASE
    return hipErrorInvalidValue;
}

template <int M>
static hipError_t launch_refine_wg_m(const RefineArgs& g)
{
    const int W = g.L / kWave;
    const size_t lds = wg_core_lds_bytes<float, M>(W) + (size_t)2 * W * sizeof(double);
    if (lds > 64 * 1024) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL((refine_wg_kernel<M>), dim3((unsigned)g.N), dim3(kWave * W), lds, g.stream,
                       g.a, g.b, g.c, g.d, g.x, g.n, g.rest, g.iters, g.vec);
    return hipGetLastError();
}

hipError_t launch_refine(const RefineArgs& g)
{
    if (g.N <= 0) return hipSuccess;
    if (g.L <= kWave) {
        const int R = g.rest > g.L ? g.rest / g.L : 1;
        const int64_t threads = g.N * (int64_t)g.L;
        const int64_t blocks = (threads + 255) / 256;
        if (blocks > 0x7fffffffLL) return hipErrorInvalidConfiguration;
        dim3 grid((unsigned)blocks), blk(256);
        switch (g.M) {
            case 1: return launch_refine_wave_m<1>(g, R, grid, blk);
            case 2: return launch_refine_wave_m<2>(g, R, grid, blk);
            case 4: return launch_refine_wave_m<4>(g, R, grid, blk);
            case 8: return launch_refine_wave_m<8>(g, R, grid, blk);
            case 16: return launch_refine_wave_m<16>(g, R, grid, blk);
            default: return hipErrorInvalidValue;
        }
    }
    // multi-wave systems: at most 8 unknowns per lane (make_refine_plan refuses
    // more; 16 per lane would not fit the registers next to the FP64 rows)
    if (g.L % kWave != 0 || g.L / kWave > 16 || g.N > 0x7fffffffLL) return hipErrorInvalidValue;
    switch (g.M) {
        case 2: return launch_refine_wg_m<2>(g);
        case 4: return launch_refine_wg_m<4>(g);
        case 8: return launch_refine_wg_m<8>(g);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace trid
