// trid_launch.hpp -- host-side launchers of the gfx950 kernels (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace trid {

// Register-resident CR / CR-PCR / PCR (kern_reg.hip).
template <typename T>
struct RegArgs {
    const T* a;
    const T* b;
    const T* c;
    const T* d;
    T* x;
    int64_t n;       // unknowns per system
    int64_t N;       // systems
    int M;           // unknowns per lane (1, 2, 4, 8, 16)
    int L;           // lanes per system (power of two, 2..64), L*M >= n
    int rest;        // size of the PCR system (power of two, 2..L*M)
    int vec;         // 1: dense 16-byte vector path (n == L*M, aligned)
    int io;          // dense-path I/O mode: 0 per-lane non-temporal, 1 per-lane,
                     // 2 LDS-staged coalesced, 3 LDS-staged non-temporal
    hipStream_t stream;
};

template <typename T>
hipError_t launch_reg_crpcr(const RegArgs<T>& g);

// Seeded synthetic batch (kern_util.hip).
template <typename T>
hipError_t launch_generate(int kind, uint64_t seed, T* a, T* b, T* c, T* d, int64_t n,
                           int64_t N, hipStream_t stream);

}  // namespace trid
