// kern_adi.hip -- Peaceman-Rachford ADI steps of the 2D heat equation on an
// n x n grid (row-major, homogeneous Dirichlet boundary), the natural caller
// of the batched tridiagonal solve. No counterpart in the reference.
//
// One step with parameter r (the tridiagonal operator of each sweep is
// tridiag(-r, 1+2r, -r), the same for every row and every column):
//   (I - r dxx) u* = (I + r dyy) u        row sweep:    n systems, contiguous
//   (I - r dyy) u' = (I + r dxx) u*       column sweep: n systems, stride n
//
// * the row sweep's right-hand side is one elementwise pass (adi_rhs_rows);
//   the solve is the register CR-PCR kernel in its constant-coefficient mode,
//   which reads only d and writes x (16 bytes per unknown in FP64);
// * the column sweep runs one lane per column over the interleaved layout:
//   at every row the 64 lanes of a wave touch 64 consecutive doubles, so both
//   passes are coalesced. Its right-hand side is formed on the fly, and the
//   Thomas factors (shared by every column) are precomputed once per r:
//   m_i = lambda1 (1 - q^(i+2)) / (1 - q^(i+1)), q = lambda2 / lambda1, the
//   closed form of the pivots of tridiag(-r, 1+2r, -r) (lambda1,2 the roots of
//   t^2 - (1+2r) t + r^2), so all n pivots are computed independently.
//   The forward pass writes d' to a scratch grid and the backward pass
//   (a separate launch, so no wave reads a neighbour's column while another
//   overwrites it) writes u'.
#include "trid_device.hpp"
#include "trid_launch.hpp"

namespace trid {

// D = u + r * (u[i-1][j] - 2 u[i][j] + u[i+1][j]), zero outside the grid.
__global__ __launch_bounds__(256) void adi_rhs_rows_kernel(const double* __restrict__ U,
                                                           double* __restrict__ D, int64_t n,
                                                           double r)
{
    const int64_t total = n * n;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {
        const int64_t i = g / n;
        const double u = U[g];
        const double up = i > 0 ? U[g - n] : 0.0;
        const double dn = i < n - 1 ? U[g + n] : 0.0;
        D[g] = u + r * (up - 2.0 * u + dn);
    }
}

// Pivot factors of tridiag(-r, 1+2r, -r): inv_m[i] = 1/m_i, cp[i] = -r/m_i.
__global__ __launch_bounds__(256) void adi_factors_kernel(double* __restrict__ cp,
                                                          double* __restrict__ inv_m, int64_t n,
                                                          double r)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double b = 1.0 + 2.0 * r;
    const double s = sqrt(1.0 + 4.0 * r);  // sqrt(b^2 - 4 r^2)
    const double l1 = 0.5 * (b + s), l2 = 0.5 * (b - s);
    const double q = l2 / l1;
    const double qi1 = pow(q, (double)(i + 1));
    const double m = l1 * (1.0 - qi1 * q) / (1.0 - qi1);
    inv_m[i] = 1.0 / m;
    cp[i] = (i < n - 1) ? -r / m : 0.0;
}

// Column sweep, forward: lane j walks column j of u*, forming
// d_i = u*[i][j] + r (u*[i][j-1] - 2 u*[i][j] + u*[i][j+1]) and
// dp_i = (d_i + r dp_{i-1}) / m_i (a = -r). dp goes to P.
__global__ __launch_bounds__(256) void adi_cols_forward_kernel(const double* __restrict__ U,
                                                               double* __restrict__ P,
                                                               const double* __restrict__ inv_m,
                                                               int64_t n, double r)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    constexpr int K = 8;  // rows read ahead
    double dp = 0.0;
    for (int64_t i0 = 0; i0 < n; i0 += K) {
        double d[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int64_t i = i0 + k;
            if (i < n) {
                const int64_t g = i * n + j;
                const double u = U[g];
                const double ul = j > 0 ? U[g - 1] : 0.0;
                const double ur = j < n - 1 ? U[g + 1] : 0.0;
                d[k] = u + r * (ul - 2.0 * u + ur);
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int64_t i = i0 + k;
            if (i < n) {
                dp = (d[k] + r * dp) * inv_m[i];
                P[i * n + j] = dp;
            }
        }
    }
}

// Column sweep, backward: x_i = dp_i - cp_i x_{i+1}, written to X.
__global__ __launch_bounds__(256) void adi_cols_backward_kernel(const double* __restrict__ P,
                                                                double* __restrict__ X,
                                                                const double* __restrict__ cp,
                                                                int64_t n)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    constexpr int K = 8;
    double x = 0.0;
    for (int64_t i0 = n - 1; i0 >= 0; i0 -= K) {
        double p[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int64_t i = i0 - k;
            if (i >= 0) p[k] = P[i * n + j];
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int64_t i = i0 - k;
            if (i >= 0) {
                x = p[k] - cp[i] * x;
                X[i * n + j] = x;
            }
        }
    }
}

hipError_t launch_adi_step(const AdiArgs& g)
{
    const int64_t n = g.n;
    const int64_t total = n * n;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    const unsigned colblocks = (unsigned)((n + 255) / 256);
    hipError_t e;
    // factors of this r (shared by every column)
    hipLaunchKernelGGL(adi_factors_kernel, dim3(colblocks), dim3(256), 0, g.stream, g.cp,
                       g.inv_m, n, g.r);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // row sweep: rhs into T, solve T -> U (u*)
    hipLaunchKernelGGL(adi_rhs_rows_kernel, dim3((unsigned)blocks), dim3(256), 0, g.stream, g.u,
                       g.t, n, g.r);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    RegArgs<double> ra{};
    ra.a = ra.b = ra.c = nullptr;
    ra.d = g.t;
    ra.x = g.u;
    ra.n = n;
    ra.N = n;
    ra.M = g.M;
    ra.L = g.L;
    ra.rest = g.rest;
    ra.vec = g.vec;
    ra.io = 3;
    ra.stream = g.stream;
    ra.cc = 1;
    ra.a0 = -g.r;
    ra.b0 = 1.0 + 2.0 * g.r;
    ra.c0 = -g.r;
    e = ra.L <= kWave ? launch_reg_crpcr<double>(ra) : launch_reg_mw<double>(ra);
    if (e != hipSuccess) return e;
    // column sweep: forward (U -> T as d'), backward (T -> U)
    hipLaunchKernelGGL(adi_cols_forward_kernel, dim3(colblocks), dim3(256), 0, g.stream, g.u,
                       g.t, g.inv_m, n, g.r);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(adi_cols_backward_kernel, dim3(colblocks), dim3(256), 0, g.stream, g.t,
                       g.u, g.cp, n);
    return hipGetLastError();
}

}  // namespace trid
