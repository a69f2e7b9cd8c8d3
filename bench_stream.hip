// bench_stream.hip -- measurement aid for bench.py (not part of libtrid.so):
// the speed of light of the solve's own traffic. One kernel reads four
// arrays and writes one, 16 bytes per lane per array with the non-temporal
// hint, one vector per lane (the fastest of the grid shapes and unroll
// depths tried), i.e. exactly the bytes of a batched solve (read a, b, c, d;
// write x) with no arithmetic in between. Built by __graft_entry__.build()
// into bench_stream.so.
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void stream_4r1w_kernel(const u32x4* __restrict__ a,
                                                          const u32x4* __restrict__ b,
                                                          const u32x4* __restrict__ c,
                                                          const u32x4* __restrict__ d,
                                                          u32x4* __restrict__ x, int64_t n16)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n16) return;
    const u32x4 va = __builtin_nontemporal_load(a + i);
    const u32x4 vb = __builtin_nontemporal_load(b + i);
    const u32x4 vc = __builtin_nontemporal_load(c + i);
    const u32x4 vd = __builtin_nontemporal_load(d + i);
    __builtin_nontemporal_store(va ^ vb ^ vc ^ vd, x + i);
}

// n16: 16-byte vectors per array. All pointers 16-byte aligned device memory.
// Returns a hipError_t value.
extern "C" int bench_stream_4r1w(const void* a, const void* b, const void* c, const void* d,
                                 void* x, int64_t n16, void* stream)
{
    if (n16 <= 0) return 0;
    const int64_t blocks = (n16 + 255) / 256;
    if (blocks > 0x7fffffffLL) return (int)hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(stream_4r1w_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       static_cast<hipStream_t>(stream), static_cast<const u32x4*>(a),
                       static_cast<const u32x4*>(b), static_cast<const u32x4*>(c),
                       static_cast<const u32x4*>(d), static_cast<u32x4*>(x), n16);
    return (int)hipGetLastError();
}
