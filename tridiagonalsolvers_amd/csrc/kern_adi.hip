// kern_adi.hip -- Peaceman-Rachford ADI steps of the 2D heat equation on an
// n x n grid (row-major, homogeneous Dirichlet boundary), the natural caller
// of the batched tridiagonal solve. No counterpart in the reference.
//
// One step with parameter r (the tridiagonal operator of each sweep is
// tridiag(-r, 1+2r, -r), the same for every row and every column):
//   (I - r dxx) u* = (I + r dyy) u        row sweep:    n systems, contiguous
//   (I - r dyy) u' = (I + r dxx) u*       column sweep: n systems, stride n
//
// Both sweeps solve systems with the same operator, so its LU factors are
// resident for the whole call (adi_factor_kernel, three values per row) and
// every system is solved by the two substitutions of the factored operator
// (adi_lu_kernel: d read, x written, 16 bytes per unknown in FP64).
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
#include "reg_io.hpp"
#include "trid_device.hpp"
#include "trid_launch.hpp"

namespace trid {

namespace {
constexpr int kTile = 64;
constexpr int kAdiM = 16;                  // unknowns per lane in the solve
constexpr int kAdiRows = kWave * kAdiM;    // rows per wave
}

int64_t adi_npad(int64_t n) { return (n + kAdiRows - 1) / kAdiRows * kAdiRows; }

// LU factors of tridiag(-r, 1+2r, -r) of order n, padded to npad rows. With
// the pivots u_0 = 1 + 2r, u_i = 1 + 2r - r^2 / u_{i-1} and q_i = 1 / u_i,
// the two substitutions are
//   y_i = d_i + r q_{i-1} y_{i-1}     (nothing before the first row)
//   x_i = q_i (y_i + r x_{i+1})       (nothing after the last row),
// so the table holds q only. With l1 > l2 the roots of l^2 - (1+2r) l + r^2
// (l1 l2 = r^2) and p = l2 / l1 < 1, the pivots have the closed form
// u_i = l1 (1 - p^(i+2)) / (1 - p^(i+1)) (induction on the recurrence),
// evaluated here per row with expm1 so that every row is independent.
// Rows past n: q = 0. Then x = 0 there whatever y is, so x_{n-1} =
// q_{n-1} y_{n-1} as it must be, and the kernel needs no mask (with q = 0
// taken before the system's first row as well).
__global__ __launch_bounds__(256) void adi_factor_kernel(double* __restrict__ q, int64_t n,
                                                         int64_t npad, double r)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad) return;
    const double l1 = 0.5 * (1.0 + 2.0 * r + sqrt(1.0 + 4.0 * r));
    const double lg = 2.0 * log(r / l1);  // log(l2 / l1)
    q[i] = i < n ? 1.0 / (l1 * (expm1((double)(i + 2) * lg) / expm1((double)(i + 1) * lg))) : 0.0;
}

// One system (a row of the grid in D, n unknowns) per workgroup of W waves,
// lane l of wave w holding rows 16 (64 w + l) .. +15; d and the factors q come
// in through the wave-coalesced staged loads. Both substitutions are linear
// recurrences: each lane composes its 16 rows into one affine map (y_out =
// A y_in + B forward, x_first = G x_in + H backward), a Kogge-Stone scan
// composes the maps across the lanes (DPP at distance 1, ds_bpermute
// beyond), the wave maps cross the waves through LDS, and each lane then
// runs its rows from its own input. About 200 FP64 operations per lane and
// no division, against about 500 for the CR-PCR solve of the same system.
template <int MAXT>
__global__ __launch_bounds__(MAXT) void adi_lu_kernel(const double* __restrict__ D,
                                                      double* __restrict__ X, int64_t n,
                                                      int64_t npad, const double* __restrict__ Q,
                                                      double r, int vec)
{
    const int W = blockDim.x / kWave;
    const int w = threadIdx.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t base = (int64_t)blockIdx.x * n;
    const int64_t row0 = (int64_t)(w * kWave + lane) * kAdiM;
    const int64_t wave_base = base + (int64_t)w * kAdiRows;
    extern __shared__ __align__(16) char trid_smem[];
    char* lds = trid_smem + w * Stage<double, kAdiM>::kWaveBytes;
    __shared__ double wmap[2][16][2];  // per wave: forward (A, B), backward (G, H)
    lds_poison_wg(wmap, (unsigned)sizeof(wmap));

    // q of the row before the lane's first (0 for the system's first row)
    const double q_prev_wave = (lane == 0 && row0 > 0) ? Q[row0 - 1] : 0.0;
    double y[kAdiM], q[kAdiM];
    {
        Eq<double> e[kAdiM];
        load_staged<double, kAdiM, false>(Q, Q, Q, Q, (int64_t)w * kAdiRows, npad, lane, lds, e,
                                          ConstCoef<double>{true, 0.0, 0.0, 0.0});
#pragma unroll
        for (int k = 0; k < kAdiM; ++k) q[k] = e[k].d;
    }
    if (vec) {
        Eq<double> e[kAdiM];
        load_staged<double, kAdiM, true>(D, D, D, D, wave_base, base + n, lane, lds, e,
                                         ConstCoef<double>{true, 0.0, 0.0, 0.0});
#pragma unroll
        for (int k = 0; k < kAdiM; ++k) y[k] = e[k].d;
    } else {
#pragma unroll
        for (int k = 0; k < kAdiM; ++k) y[k] = row0 + k < n ? D[base + row0 + k] : 0.0;
    }
    // (the DPP move runs on every lane: inside a branch that lane 0 skips,
    // lane 1 would read the inactive lane 0 and get 0)
    const double q_left = lane_step<-1>(q[kAdiM - 1]);
    const double q_prev = lane == 0 ? q_prev_wave : q_left;

    // ---- forward: y_i = d_i + r q_{i-1} y_{i-1} ------------------------------
    {
        double A = 1.0, B = 0.0;
#pragma unroll
        for (int k = 0; k < kAdiM; ++k) {
            const double f = r * (k == 0 ? q_prev : q[k > 0 ? k - 1 : 0]);
            B = fma(f, B, y[k]);
            A *= f;
        }
        {  // inclusive scan over the lanes: lane l maps the wave's input to the end of its rows
            const double Ap = lane_step<-1>(A), Bp = lane_step<-1>(B);
            if (lane >= 1) {
                B = fma(A, Bp, B);
                A *= Ap;
            }
        }
#pragma unroll
        for (int sh = 2; sh < kWave; sh <<= 1) {
            const double Ap = lane_read(A, (lane - sh) & (kWave - 1));
            const double Bp = lane_read(B, (lane - sh) & (kWave - 1));
            if (lane >= sh) {
                B = fma(A, Bp, B);
                A *= Ap;
            }
        }
        if (lane == kWave - 1) {
            wmap[0][w][0] = A;
            wmap[0][w][1] = B;
        }
        __syncthreads();
        double yin = 0.0;
        for (int v = 0; v < w; ++v) yin = fma(wmap[0][v][0], yin, wmap[0][v][1]);
        const double Ap = lane_step<-1>(A), Bp = lane_step<-1>(B);
        double t = lane == 0 ? yin : fma(Ap, yin, Bp);
        double qp = q_prev;
#pragma unroll
        for (int k = 0; k < kAdiM; ++k) {
            t = fma(r, qp * t, y[k]);
            y[k] = t;
            qp = q[k];
        }
    }

    // ---- backward: x_i = q_i (y_i + r x_{i+1}) -------------------------------
    {
        double G = 1.0, H = 0.0;
#pragma unroll
        for (int k = kAdiM - 1; k >= 0; --k) {
            H = q[k] * fma(r, H, y[k]);
            G *= r * q[k];
        }
        {  // suffix scan: lane l maps the wave's right input to its first row
            const double Gn = lane_step<1>(G), Hn = lane_step<1>(H);
            if (lane + 1 < kWave) {
                H = fma(G, Hn, H);
                G *= Gn;
            }
        }
#pragma unroll
        for (int sh = 2; sh < kWave; sh <<= 1) {
            const double Gn = lane_read(G, (lane + sh) & (kWave - 1));
            const double Hn = lane_read(H, (lane + sh) & (kWave - 1));
            if (lane + sh < kWave) {
                H = fma(G, Hn, H);
                G *= Gn;
            }
        }
        if (lane == 0) {
            wmap[1][w][0] = G;
            wmap[1][w][1] = H;
        }
        __syncthreads();
        double xin = 0.0;
        for (int v = W - 1; v > w; --v) xin = fma(wmap[1][v][0], xin, wmap[1][v][1]);
        const double Gn = lane_step<1>(G), Hn = lane_step<1>(H);
        double x = lane == kWave - 1 ? xin : fma(Gn, xin, Hn);
#pragma unroll
        for (int k = kAdiM - 1; k >= 0; --k) {
            x = q[k] * fma(r, x, y[k]);
            y[k] = x;
        }
    }

    if (vec) {
        store_staged<double, kAdiM, true>(X, wave_base, base + n, lane, lds, y);
    } else {
#pragma unroll
        for (int k = 0; k < kAdiM; ++k)
            if (row0 + k < n) X[base + row0 + k] = y[k];
    }
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
    lds_poison_wg(tile, (unsigned)sizeof(tile));
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
    lds_poison_wg(tile, (unsigned)sizeof(tile));
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
    // the operator's factors, once for the call
    const int64_t npad = adi_npad(n);
    hipLaunchKernelGGL(adi_factor_kernel, dim3((unsigned)((npad + 255) / 256)), dim3(256), 0,
                       g.stream, g.fac, n, npad, g.r);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const int W = (int)(npad / kAdiRows);
    const size_t lu_lds = (size_t)W * Stage<double, kAdiM>::kWaveBytes;
    if (W > 16) return hipErrorInvalidConfiguration;
    // up to 8 waves: 512-thread bound (registers for 2-3 workgroups per CU);
    // 9..16 waves: the 1024-thread bound
    const bool wide = W > 8;
    if (wide && lu_lds > 64 * 1024) {
        e = hipFuncSetAttribute((const void*)adi_lu_kernel<1024>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lu_lds);
        if (e != hipSuccess) return e;
    }
    auto solve = [&](const double* d, double* x) -> hipError_t {
        if (wide)
            hipLaunchKernelGGL((adi_lu_kernel<1024>), dim3((unsigned)n), dim3(W * kWave), lu_lds,
                               g.stream, d, x, n, npad, g.fac, g.r, tvec);
        else
            hipLaunchKernelGGL((adi_lu_kernel<512>), dim3((unsigned)n), dim3(W * kWave), lu_lds,
                               g.stream, d, x, n, npad, g.fac, g.r, tvec);
        return hipGetLastError();
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
