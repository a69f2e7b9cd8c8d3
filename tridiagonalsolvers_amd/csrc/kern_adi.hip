// kern_adi.hip -- Peaceman-Rachford ADI steps of the 2D heat equation on an
// n x n grid (row-major, homogeneous Dirichlet boundary), the natural caller
// of the batched tridiagonal solve. No counterpart in the reference.
//
// One step with parameter r (the tridiagonal operator of each sweep is
// tridiag(-r, 1+2r, -r), the same for every row and every column):
//   (I - r dxx) u* = (I + r dyy) u        row sweep:    n systems, contiguous
//   (I - r dyy) u' = (I + r dxx) u*       column sweep: n systems, stride n
//
// Both sweeps run on the register CR-PCR kernels in their constant-
// coefficient mode (only d is read, x written: 16 bytes per unknown in FP64).
// * row sweep: the right-hand side is one elementwise pass (adi_rhs_rows),
//   then the row solve;
// * column sweep: the right-hand side is formed and transposed in one tiled
//   pass (adi_rhs_cols_t: 64 x 64 tiles through LDS, reads and writes
//   coalesced), the columns are solved as the rows of the transposed grid,
//   and a tiled transpose writes u'. A direct column solve (one lane per
//   column, n/64 waves) leaves the GPU mostly idle; the transposes cost two
//   streaming passes and keep every pass at full width.
// * several steps in one call: u' of a step stays transposed, and the next
//   row rhs is formed from it by the same transposing stencil kernel (the
//   stencil along the transposed grid's rows is the stencil along u's
//   columns); only the last step transposes u' back into u: 4 passes a step
//   instead of 5.
#include "trid_device.hpp"
#include "trid_launch.hpp"

namespace trid {

namespace {
constexpr int kTile = 64;
}

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

// Dt[j][i] = u[i][j] + r * (u[i][j-1] - 2 u[i][j] + u[i][j+1]) for one 64 x 64
// tile: rows i0.., columns j0-1 .. j0+64 are read (coalesced along j) into
// LDS, the tile is written transposed (coalesced along i). Row pitch 67
// doubles spreads the column reads tile[tx][c] over the banks.
__global__ __launch_bounds__(256) void adi_rhs_cols_t_kernel(const double* __restrict__ U,
                                                             double* __restrict__ Dt, int64_t n,
                                                             double r)
{
    __shared__ double tile[kTile][kTile + 3];
    const int tx = threadIdx.x & (kTile - 1);
    const int ty = threadIdx.x / kTile;  // 0..3
    const int64_t i0 = (int64_t)blockIdx.y * kTile;
    const int64_t j0 = (int64_t)blockIdx.x * kTile;
    // the 64 columns of the tile as one aligned 512-byte row per wave; the
    // two halo columns j0-1 and j0+64 by two lanes
    for (int row = ty; row < kTile; row += 4) {
        const int64_t i = i0 + row;
        const int64_t j = j0 + tx;
        tile[row][tx + 1] = (i < n && j < n) ? U[i * n + j] : 0.0;
        if (tx < 2) {
            const int64_t jh = tx == 0 ? j0 - 1 : j0 + kTile;
            tile[row][tx == 0 ? 0 : kTile + 1] =
                (i < n && jh >= 0 && jh < n) ? U[i * n + jh] : 0.0;
        }
    }
    __syncthreads();
    const int64_t i = i0 + tx;
    for (int jj = ty; jj < kTile; jj += 4) {
        const int64_t j = j0 + jj;
        if (i < n && j < n) {
            const double u = tile[tx][jj + 1];
            Dt[j * n + i] = u + r * (tile[tx][jj] - 2.0 * u + tile[tx][jj + 2]);
        }
    }
}

// U[i][j] = Xt[j][i], 64 x 64 tiles through LDS (pitch 65 doubles).
__global__ __launch_bounds__(256) void adi_transpose_kernel(const double* __restrict__ Xt,
                                                            double* __restrict__ U, int64_t n)
{
    __shared__ double tile[kTile][kTile + 1];
    const int tx = threadIdx.x & (kTile - 1);
    const int ty = threadIdx.x / kTile;
    const int64_t j0 = (int64_t)blockIdx.y * kTile;  // rows of Xt
    const int64_t i0 = (int64_t)blockIdx.x * kTile;  // columns of Xt
    for (int jj = ty; jj < kTile; jj += 4) {
        const int64_t j = j0 + jj, i = i0 + tx;
        if (j < n && i < n) tile[jj][tx] = Xt[j * n + i];
    }
    __syncthreads();
    for (int ii = ty; ii < kTile; ii += 4) {
        const int64_t i = i0 + ii, j = j0 + tx;
        if (i < n && j < n) U[i * n + j] = tile[tx][ii];
    }
}

hipError_t launch_adi_steps(const AdiArgs& g, int steps)
{
    const int64_t n = g.n;
    const int64_t total = n * n;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    const int64_t tiles = (n + kTile - 1) / kTile;
    if (tiles > 65535) return hipErrorInvalidConfiguration;
    hipError_t e;
    RegArgs<double> ra{};
    ra.a = ra.b = ra.c = nullptr;
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
    auto solve = [&](const double* d, double* x) -> hipError_t {
        ra.d = d;
        ra.x = x;
        return ra.L <= kWave ? launch_reg_crpcr<double>(ra) : launch_reg_mw<double>(ra);
    };
    const dim3 tgrid((unsigned)tiles, (unsigned)tiles);
    for (int k = 0; k < steps; ++k) {
        // row sweep: rhs into T, solve T -> U (u*). After a step whose u' was
        // left transposed in T2, the same tiled kernel forms the row rhs from
        // T2 (transposing it back on the way), saving a transpose and a pass.
        if (k == 0)
            hipLaunchKernelGGL(adi_rhs_rows_kernel, dim3((unsigned)blocks), dim3(256), 0,
                               g.stream, g.u, g.t, n, g.r);
        else
            hipLaunchKernelGGL(adi_rhs_cols_t_kernel, tgrid, dim3(256), 0, g.stream, g.t2, g.t,
                               n, g.r);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = solve(g.t, g.u)) != hipSuccess) return e;
        // column sweep: rhs transposed into T, solve T -> T2 (u' transposed)
        hipLaunchKernelGGL(adi_rhs_cols_t_kernel, tgrid, dim3(256), 0, g.stream, g.u, g.t, n,
                           g.r);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = solve(g.t, g.t2)) != hipSuccess) return e;
    }
    // the grid itself after the last step: transpose T2 -> U
    if (steps > 0) {
        hipLaunchKernelGGL(adi_transpose_kernel, tgrid, dim3(256), 0, g.stream, g.t2, g.u, n);
        return hipGetLastError();
    }
    return hipSuccess;
}

}  // namespace trid
