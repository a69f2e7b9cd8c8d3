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

typedef double adi_d2 __attribute__((ext_vector_type(2)));

// D = u + r * (u[i-1][j] - 2 u[i][j] + u[i+1][j]), zero outside the grid.
// Two consecutive points per thread with 16-byte accesses when n is even
// (vec; a pair then never straddles two rows), one otherwise.
__global__ __launch_bounds__(256) void adi_rhs_rows_kernel(const double* __restrict__ U,
                                                           double* __restrict__ D, int64_t n,
                                                           double r, int vec)
{
    const int64_t total = n * n;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    if (vec) {
        for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; 2 * p < total;
             p += stride) {
            const int64_t g = 2 * p;
            const int64_t i = g / n;
            const adi_d2 u = *reinterpret_cast<const adi_d2*>(U + g);
            const adi_d2 up = i > 0 ? *reinterpret_cast<const adi_d2*>(U + g - n) : adi_d2{0, 0};
            const adi_d2 dn =
                i < n - 1 ? *reinterpret_cast<const adi_d2*>(U + g + n) : adi_d2{0, 0};
            adi_d2 o;
            o.x = u.x + r * (up.x - 2.0 * u.x + dn.x);
            o.y = u.y + r * (up.y - 2.0 * u.y + dn.y);
            *reinterpret_cast<adi_d2*>(D + g) = o;
        }
        return;
    }
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {
        const int64_t i = g / n;
        const double u = U[g];
        const double up = i > 0 ? U[g - n] : 0.0;
        const double dn = i < n - 1 ? U[g + n] : 0.0;
        D[g] = u + r * (up - 2.0 * u + dn);
    }
}

// Tile (ty, tx) of block b of a one-dimensional grid of tiles x tiles blocks.
// Blocks are dealt to the 8 XCDs round-robin; when the tile count is a
// multiple of 8, XCD k takes the k-th eighth of the tiles in row order, so
// neighbouring tiles (which share the lines of their halo columns and of
// their rows) run together on one XCD's L2.
__device__ __forceinline__ void adi_tile(int64_t b, int64_t tiles, int64_t& ty, int64_t& tx)
{
    const int64_t T = tiles * tiles;
    const int64_t L = (T % 8 == 0) ? (b % 8) * (T / 8) + b / 8 : b;
    ty = L / tiles;
    tx = L - ty * tiles;
}

// Elements j, j + 1 of row i (zero outside the grid); 16-byte load when the
// rows are 16-byte aligned (vec: n even) and both are inside.
__device__ __forceinline__ void adi_load2(const double* __restrict__ U, int64_t n, int64_t i,
                                          int64_t j, bool vec, double& v0, double& v1)
{
    v0 = 0.0;
    v1 = 0.0;
    if (i >= n) return;
    if (vec && j + 1 < n) {
        const adi_d2 q = *reinterpret_cast<const adi_d2*>(U + i * n + j);
        v0 = q.x;
        v1 = q.y;
    } else {
        if (j < n) v0 = U[i * n + j];
        if (j + 1 < n) v1 = U[i * n + j + 1];
    }
}

__device__ __forceinline__ void adi_store2(double* __restrict__ D, int64_t n, int64_t j,
                                           int64_t i, bool vec, double v0, double v1)
{
    if (j >= n) return;
    if (vec && i + 1 < n) {
        adi_d2 q;
        q.x = v0;
        q.y = v1;
        *reinterpret_cast<adi_d2*>(D + j * n + i) = q;
    } else {
        if (i < n) D[j * n + i] = v0;
        if (i + 1 < n) D[j * n + i + 1] = v1;
    }
}

// Dt[j][i] = u[i][j] + r * (u[i][j-1] - 2 u[i][j] + u[i][j+1]) for one 64 x 64
// tile: rows i0.., columns j0-1 .. j0+64 are read into LDS (32 lanes per
// row, two elements each: a wave reads two 512-byte rows per instruction; the
// two halo columns by 128 single loads), the tile is written transposed the
// same way. Row pitch 67 doubles.
__global__ __launch_bounds__(256) void adi_rhs_cols_t_kernel(const double* __restrict__ U,
                                                             double* __restrict__ Dt, int64_t n,
                                                             double r, int vec)
{
    __shared__ double tile[kTile][kTile + 3];
    const int t = threadIdx.x;
    const int cx = (t & 31) * 2;  // column pair of the tile
    const int ry = t / 32;        // 0..7
    int64_t ty, tx;
    adi_tile(blockIdx.x, (n + kTile - 1) / kTile, ty, tx);
    const int64_t i0 = ty * kTile;
    const int64_t j0 = tx * kTile;
    for (int row = ry; row < kTile; row += 8) {
        double v0, v1;
        adi_load2(U, n, i0 + row, j0 + cx, vec != 0, v0, v1);
        tile[row][cx + 1] = v0;
        tile[row][cx + 2] = v1;
    }
    if (t < 2 * kTile) {
        const int row = t >> 1;
        const int64_t i = i0 + row;
        const int64_t jh = (t & 1) ? j0 + kTile : j0 - 1;
        tile[row][(t & 1) ? kTile + 1 : 0] = (i < n && jh >= 0 && jh < n) ? U[i * n + jh] : 0.0;
    }
    __syncthreads();
    for (int jj = ry; jj < kTile; jj += 8) {
        const double u0 = tile[cx][jj + 1];
        const double u1 = tile[cx + 1][jj + 1];
        const double r0 = u0 + r * (tile[cx][jj] - 2.0 * u0 + tile[cx][jj + 2]);
        const double r1 = u1 + r * (tile[cx + 1][jj] - 2.0 * u1 + tile[cx + 1][jj + 2]);
        adi_store2(Dt, n, j0 + jj, i0 + cx, vec != 0, r0, r1);
    }
}

// U[i][j] = Xt[j][i], 64 x 64 tiles through LDS (pitch 65 doubles), the
// same access shape as the stencil kernel.
__global__ __launch_bounds__(256) void adi_transpose_kernel(const double* __restrict__ Xt,
                                                            double* __restrict__ U, int64_t n,
                                                            int vec)
{
    __shared__ double tile[kTile][kTile + 1];
    const int t = threadIdx.x;
    const int cx = (t & 31) * 2;
    const int ry = t / 32;
    int64_t ty, tx;
    adi_tile(blockIdx.x, (n + kTile - 1) / kTile, ty, tx);
    const int64_t j0 = ty * kTile;  // rows of Xt
    const int64_t i0 = tx * kTile;  // columns of Xt
    for (int jj = ry; jj < kTile; jj += 8) {
        double v0, v1;
        adi_load2(Xt, n, j0 + jj, i0 + cx, vec != 0, v0, v1);
        tile[jj][cx] = v0;
        tile[jj][cx + 1] = v1;
    }
    __syncthreads();
    for (int ii = ry; ii < kTile; ii += 8)
        adi_store2(U, n, i0 + ii, j0 + cx, vec != 0, tile[cx][ii], tile[cx + 1][ii]);
}

hipError_t launch_adi_steps(const AdiArgs& g, int steps)
{
    const int64_t n = g.n;
    const int64_t total = n * n;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    const int64_t tiles = (n + kTile - 1) / kTile;
    if (tiles * tiles > 0x7fffffffLL) return hipErrorInvalidConfiguration;
    const int tvec = (n % 2 == 0 && (reinterpret_cast<uintptr_t>(g.u) & 15u) == 0 &&
                      (reinterpret_cast<uintptr_t>(g.t) & 15u) == 0 &&
                      (reinterpret_cast<uintptr_t>(g.t2) & 15u) == 0)
                         ? 1
                         : 0;
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
    const dim3 tgrid((unsigned)(tiles * tiles));
    for (int k = 0; k < steps; ++k) {
        // row sweep: rhs into T, solve T -> U (u*). After a step whose u' was
        // left transposed in T2, the same tiled kernel forms the row rhs from
        // T2 (transposing it back on the way), saving a transpose and a pass.
        if (k == 0)
            hipLaunchKernelGGL(adi_rhs_rows_kernel, dim3((unsigned)blocks), dim3(256), 0,
                               g.stream, g.u, g.t, n, g.r, tvec);
        else
            hipLaunchKernelGGL(adi_rhs_cols_t_kernel, tgrid, dim3(256), 0, g.stream, g.t2, g.t,
                               n, g.r, tvec);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = solve(g.t, g.u)) != hipSuccess) return e;
        // column sweep: rhs transposed into T, solve T -> T2 (u' transposed)
        hipLaunchKernelGGL(adi_rhs_cols_t_kernel, tgrid, dim3(256), 0, g.stream, g.u, g.t, n,
                           g.r, tvec);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = solve(g.t, g.t2)) != hipSuccess) return e;
    }
    // the grid itself after the last step: transpose T2 -> U
    if (steps > 0) {
        hipLaunchKernelGGL(adi_transpose_kernel, tgrid, dim3(256), 0, g.stream, g.t2, g.u, n,
                           tvec);
        return hipGetLastError();
    }
    return hipSuccess;
}

}  // namespace trid
