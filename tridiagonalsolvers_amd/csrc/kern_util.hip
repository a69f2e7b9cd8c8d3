// kern_util.hip -- seeded synthetic batches generated on the device.
//
// Counter-based: element g of stream k (0=a 1=b 2=c 3=d) is
//   u = (splitmix64(key_k ^ g) >> 11) * 2^-53,  key_k = splitmix64(seed ^ splitmix64(k + 1))
// so any element can be produced independently and the values do not depend
// on the launch shape. splitmix64 is the output function of Steele, Lea and
// Flood's SplitMix (constants as published by Vigna). Floating-point
// contraction is off so the double arithmetic is exactly the sequence
// written here; each value is rounded once to the element type.
//
// Kinds (the reference generator, tridiagonal_gold.h:75-94, uses unseeded
// rand() and cannot be reproduced; kind 2 keeps its value ranges):
//   0  a,c ~ U[-1,1], b = |a|+|c| + U[1,2],      d ~ U[-1,1]
//   1  a,c ~ U[-1,1], b = |a|+|c| + U[0.01,0.1], d ~ U[-1,1]
//   2  b = 8+U, a = 3+U, c = 2+U, d = 5+U
//   3  not diagonally dominant, well conditioned once rows are interchanged:
//      rows 2k, 2k+1 are coupled by c_{2k} = 1+U and a_{2k+1} = 1+U, the
//      diagonals are tiny (b = 0.001 U[-1,1]) and the other couplings small
//      (0.25 U[-1,1]); the last row of an odd-sized system has b = 1+U;
//      d ~ U[-1,1]. Elimination without pivoting meets tiny pivots on it.
//   4  a, b, c, d ~ U[-1,1]: general, not dominant, may be ill-conditioned
// with a = 0 on the first and c = 0 on the last row of every system.
#include "trid_device.hpp"
#include "trid_launch.hpp"

#pragma clang fp contract(off)

namespace trid {

__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t stream_key(uint64_t seed, uint64_t stream)
{
    return splitmix64(seed ^ splitmix64(stream + 1));
}

__device__ __forceinline__ double unit(uint64_t key, uint64_t g)
{
    return (double)(splitmix64(key ^ g) >> 11) * 0x1.0p-53;
}

template <typename T>
__global__ __launch_bounds__(256) void generate_kernel(int kind, uint64_t ka, uint64_t kb,
                                                       uint64_t kc, uint64_t kd, T* A, T* B,
                                                       T* C, T* D, int64_t n, int64_t total)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {
        const int64_t j = g % n;
        const double ua = unit(ka, (uint64_t)g), ub = unit(kb, (uint64_t)g);
        const double uc = unit(kc, (uint64_t)g), ud = unit(kd, (uint64_t)g);
        T av, bv, cv, dv;
        if (kind == 2) {
            av = (T)(3.0 + ua);
            bv = (T)(8.0 + ub);
            cv = (T)(2.0 + uc);
            dv = (T)(5.0 + ud);
            if (j == 0) av = T(0);
            if (j == n - 1) cv = T(0);
        } else if (kind == 3) {
            const bool even = (j & 1) == 0;
            av = (T)(even ? 0.25 * (2.0 * ua - 1.0) : 1.0 + ua);
            bv = (T)((even && j == n - 1) ? 1.0 + ub : 0.001 * (2.0 * ub - 1.0));
            cv = (T)(even ? 1.0 + uc : 0.25 * (2.0 * uc - 1.0));
            dv = (T)(2.0 * ud - 1.0);
            if (j == 0) av = T(0);
            if (j == n - 1) cv = T(0);
        } else if (kind == 4) {
            av = (T)(2.0 * ua - 1.0);
            bv = (T)(2.0 * ub - 1.0);
            cv = (T)(2.0 * uc - 1.0);
            dv = (T)(2.0 * ud - 1.0);
            if (j == 0) av = T(0);
            if (j == n - 1) cv = T(0);
        } else {
            av = (T)(2.0 * ua - 1.0);
            cv = (T)(2.0 * uc - 1.0);
            dv = (T)(2.0 * ud - 1.0);
            if (j == 0) av = T(0);
            if (j == n - 1) cv = T(0);
            const double margin = (kind == 0) ? (1.0 + ub) : (0.01 + 0.09 * ub);
            bv = (T)((fabs((double)av) + fabs((double)cv)) + margin);
        }
        A[g] = av;
        B[g] = bv;
        C[g] = cv;
        D[g] = dv;
    }
}

template <typename T>
hipError_t launch_generate(int kind, uint64_t seed, T* a, T* b, T* c, T* d, int64_t n,
                           int64_t N, hipStream_t stream)
{
    const int64_t total = n * N;
    if (total <= 0) return hipSuccess;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL((generate_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, stream, kind,
                       stream_key(seed, 0), stream_key(seed, 1), stream_key(seed, 2),
                       stream_key(seed, 3), a, b, c, d, n, total);
    return hipGetLastError();
}

// Relative residual of every system: max_i |d - T x|_i / max_i |d_i|, FP64,
// one wave per system (a[first] and c[last] ignored).
__global__ __launch_bounds__(256) void residual_kernel(const double* __restrict__ A,
                                                      const double* __restrict__ B,
                                                      const double* __restrict__ C,
                                                      const double* __restrict__ D,
                                                      const double* __restrict__ X,
                                                      const int64_t* __restrict__ offsets,
                                                      int64_t N, double* __restrict__ out)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    if (s >= N) return;  // whole wave leaves together
    const int64_t o = offsets[s], n = offsets[s + 1] - o;
    double rmax = 0.0, dmax = 0.0;
    for (int64_t j = lane; j < n; j += kWave) {
        double r = D[o + j] - B[o + j] * X[o + j];
        if (j > 0) r -= A[o + j] * X[o + j - 1];
        if (j < n - 1) r -= C[o + j] * X[o + j + 1];
        rmax = (r != r) ? r : fmax(rmax, fabs(r));
        dmax = fmax(dmax, fabs(D[o + j]));
    }
    for (int k = kWave / 2; k >= 1; k >>= 1) {
        const double ro = __shfl_xor(rmax, k);
        rmax = (ro != ro) ? ro : fmax(rmax, ro);
        dmax = fmax(dmax, __shfl_xor(dmax, k));
    }
    if (lane == 0) out[s] = dmax > 0.0 ? rmax / dmax : rmax;
}

hipError_t launch_residual(const double* a, const double* b, const double* c, const double* d,
                           const double* x, const int64_t* offsets, int64_t N, double* out,
                           hipStream_t stream)
{
    if (N <= 0) return hipSuccess;
    const int64_t blocks = (N * kWave + 255) / 256;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(residual_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, a, b, c, d,
                       x, offsets, N, out);
    return hipGetLastError();
}

template hipError_t launch_generate<float>(int, uint64_t, float*, float*, float*, float*,
                                           int64_t, int64_t, hipStream_t);
template hipError_t launch_generate<double>(int, uint64_t, double*, double*, double*, double*,
                                            int64_t, int64_t, hipStream_t);

}  // namespace trid
