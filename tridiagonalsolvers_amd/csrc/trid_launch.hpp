// trid_launch.hpp -- host-side launchers of the gfx950 kernels (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <vector>

#include "trid.h"

namespace trid {

// Compute units of the current device, cached per device (256 when the
// query fails).
inline int device_cus()
{
    static std::atomic<int> cached[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int cus = cached[dev].load(std::memory_order_relaxed);
    if (cus == 0) {
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                hipSuccess || cus <= 0)
            cus = 256;
        cached[dev].store(cus, std::memory_order_relaxed);
    }
    return cus;
}

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
    int L;           // lanes per system (power of two, 2..1024), L*M >= n
    int rest;        // size of the PCR system (power of two, 2..L*M)
    int vec;         // 1: dense 16-byte vector path (n == L*M, aligned)
    int io;          // dense-path I/O mode: 0 per-lane non-temporal, 1 per-lane,
                     // 2 LDS-staged coalesced, 3 LDS-staged non-temporal
    hipStream_t stream;
    int64_t bs = 0;  // elements from one system's first row to the next's (0: n);
                     // with bs != n the wave-staged I/O needs L >= 64
};

template <typename T>
hipError_t launch_reg_crpcr(const RegArgs<T>& g);

// Multi-wave register-resident CR / CR-PCR (kern_reg_mw.hip): g.L = 64 * W.
template <typename T>
hipError_t launch_reg_mw(const RegArgs<T>& g);

// LDS-resident CR / CR-PCR with a pad map (kern_lds.hip).
template <typename T>
struct LdsArgs {
    const T* a;
    const T* b;
    const T* c;
    const T* d;
    T* x;
    int64_t n;
    int64_t N;
    int n_pad;   // power of two >= n
    int rest;    // PCR system size (2 = CR, n_pad = PCR)
    int pad;     // pad period in elements, 0 = none
    hipStream_t stream;
};

template <typename T>
hipError_t launch_lds_crpcr(const LdsArgs<T>& g);

// PCR to G subsystems + Thomas (kern_pcrgauss.hip).
template <typename T>
struct PcrThomasArgs {
    const T* a;
    const T* b;
    const T* c;
    const T* d;
    T* x;
    int64_t n;
    int64_t N;
    int n_pad;
    int G;          // subsystems; PCR depth log2(G)
    int pad;        // in-LDS form: pad period in elements, 0 = none
    int two_pass;   // 1: global PCR pass + strided Thomas pass through `ws`
    T* ws;          // two-pass workspace: pcrg_ws_elems(n * N) elements
    hipStream_t stream;
};
// Two-pass pcrGauss workspace: the normalized rows (3 arrays of n*N
// elements), then n*N factor pairs starting at an even element offset.
inline int64_t pcrg_pairs_offset(int64_t total) { return (3 * total + 1) & ~int64_t(1); }
inline int64_t pcrg_ws_elems(int64_t total) { return pcrg_pairs_offset(total) + 2 * total; }

template <typename T>
hipError_t launch_pcr_thomas(const PcrThomasArgs<T>& g);

// Interleaved batch (kern_pcrgauss.hip): row i of system j at i * stride + j.
template <typename T>
struct InterleavedArgs {
    const T* a;
    const T* b;
    const T* c;
    const T* d;
    T* x;
    int64_t n;       // rows per system
    int64_t N;       // systems, N <= stride
    int64_t stride;  // elements between consecutive rows of one system
    T* ws;           // workspace of n * stride elements (c' factors); unused
                     // (may be null) when interleaved_reg_takes<T>(n, N)
    hipStream_t stream;
};

template <typename T>
hipError_t launch_interleaved(const InterleavedArgs<T>& g);
// The register-core form for 256 (FP32) / 128 (FP64) < n <= 1024
// (kern_il_reg.hip); *taken is
// false when the request is outside its range and nothing was launched.
template <typename T>
hipError_t launch_interleaved_reg(const InterleavedArgs<T>& g, bool* taken);
// True when the register-core form serves (n, N); it needs no workspace.
template <typename T>
bool interleaved_reg_takes(int64_t n, int64_t N);

// Partial-pivoting solve, LAPACK ?gtsv semantics (kern_gtsv.hip).
template <typename T>
struct PivotArgs {
    const T* a;
    const T* b;
    const T* c;
    const T* d;
    T* x;            // may equal d
    int64_t n;       // rows per system
    int64_t N;       // systems
    int64_t stride;  // 0: system-major; > 0: interleaved, row i of system s at i*stride + s
    T* ws;           // checkpoints, pivot_ws_bytes<T>(n, N, stride > 0) bytes
    int32_t* info;   // per system: 0, or the 1-based row of the first zero pivot (may be null)
    hipStream_t stream;
};

template <typename T>
size_t pivot_ws_bytes(int64_t n, int64_t N, bool interleaved);
template <typename T>
hipError_t launch_pivot(const PivotArgs<T>& g);

// CR / CR-PCR by separators for systems beyond the register kernels
// (kern_part.hip): chunks of 1,024 rows, a separator system of
// part_chunks(n) unknowns per system, the chunks solved again.
template <typename T>
struct PartArgs {
    const T* a;
    const T* b;
    const T* c;
    const T* d;
    T* x;
    int64_t n;
    int64_t N;
    int64_t bs;   // elements from one system's first row to the next's (0: n)
    int method;   // TRID_CR or TRID_CRPCR (the separator system's method)
    int rest;     // PCR system size of the chunk solves: 2, 4, .., 512
    void* ws;     // part_ws_bytes<T>(method, n, N) bytes, 256-byte aligned
    hipStream_t stream;
};

int64_t part_chunks(int64_t n);
template <typename T>
size_t part_ws_bytes(int method, int64_t n, int64_t N);
template <typename T>
hipError_t launch_part(const PartArgs<T>& g);

// FP32 solve + FP64 residual refinement (kern_refine.hip).
struct RefineArgs {
    const double* a;
    const double* b;
    const double* c;
    const double* d;
    double* x;
    int64_t n;
    int64_t N;
    int M;       // unknowns per lane
    int L;       // lanes per system (<= 64: one wave; 64*W: W waves)
    int rest;    // PCR system size of the FP32 solves
    int iters;   // refinement steps
    int vec;     // 1: dense and 16-byte aligned
    hipStream_t stream;
};

hipError_t launch_refine(const RefineArgs& g);

// Ragged persistent solve (kern_ragged.hip).
struct RaggedItem {
    int32_t bucket;  // index into the bucket table
    int32_t first;   // first position in `order`
    int32_t count;   // systems in the item
    int32_t pad;
};
struct RaggedBucket {
    int32_t n_pad;     // padded size (power of two)
    int32_t per_lane;  // unknowns per lane
    int32_t lanes;     // lanes per system (64*W for workgroup items)
    int32_t rest;      // CR -> PCR switch size
};
// One system of a ragged batch, in work-queue order: first unknown and size.
struct RaggedSys {
    int64_t base;
    int64_t n;
};
struct RaggedArgs {
    const double* a;
    const double* b;
    const double* c;
    const double* d;
    double* x;
    const RaggedSys* sys;     // device, numSystems, grouped by bucket
    const RaggedItem* items;  // device: wide items, workgroup items, then wave items
    const RaggedBucket* buckets;
    int* counters;            // device, 3 queue counters (wide, 16-per-lane, 8-per-lane items)
    int n_wide_items;         // items of 8193..16384-unknown systems (the first items)
    int n_full_items;         // end of the items at 16 per lane in 8-wave workgroups
    int n_wg_items;           // end of the workgroup items (<= 8 waves after n_full_items)
    int n_wave_items;
    int n_buckets;            // <= kRaggedMaxBuckets
    int iters;                // refinement steps (the most, when tol > 0)
    int direct;               // 1: one FP64 solve per system (no FP32 stage)
    double tol;               // > 0: a system takes no more refinements once
                              // max|d - A x| <= tol max|d| (0: exactly `iters`)
};

// Grids of the four persistent kernels (wide items, 16- and 8-per-lane
// workgroup items, wave items) for a batch: every slot of the device, or
// fewer when the batch has fewer items.
struct RaggedGrids {
    int wide;
    int wg;
    int half;
    int wave;
};
RaggedGrids ragged_grids(int device, const RaggedArgs& g);
// Reset of the queue counters, then the kernels of the item classes present.
hipError_t launch_ragged(const RaggedArgs& g, const RaggedGrids& grid, hipStream_t stream);

// Host side of a ragged plan (trid_ragged.cpp): systems grouped by padded
// size, the bucket table and the work items of the persistent kernel.
constexpr int64_t kRaggedMaxSize = 16384;
constexpr int kRaggedMaxBuckets = 16;  // padded sizes 2 .. 16384
struct RaggedHostPlan {
    std::vector<int32_t> order;
    std::vector<RaggedSys> sys;     // systems in `order`
    std::vector<RaggedItem> items;  // wide items, workgroup items, then wave items
    std::vector<RaggedBucket> buckets;
    std::vector<int64_t> bucket_systems;
    int n_wide_items = 0;           // of the n_wg_items, those above 8192 unknowns
    int n_full_items = 0;           // end of the 16-per-lane 8-wave items
    int n_wg_items = 0;
    int n_wave_items = 0;
};
RaggedBucket ragged_default_bucket(int n_pad);
int ragged_check_bucket(const RaggedBucket& b);
int ragged_build_host(const int64_t* offsets, int64_t N, const trid_bucket_cfg* table,
                      int tableSize, RaggedHostPlan* out);

// Seeded synthetic batch (kern_util.hip).
template <typename T>
hipError_t launch_generate(int kind, uint64_t seed, T* a, T* b, T* c, T* d, int64_t n,
                           int64_t N, hipStream_t stream);

// Peaceman-Rachford ADI steps on an n x n FP64 grid (kern_adi.hip).
struct AdiArgs {
    double* u;      // grid, row-major, n*n (input and output), 2 <= n <= 16384
    double* t;      // scratch grid, n*n
    double* t2;     // second scratch grid, n*n
    double* fac;    // the operator's LU factors (pivot reciprocals), adi_npad(n) doubles
    int64_t n;
    double r;
    hipStream_t stream;
};

// Rows of the factor table (n rounded up to a wave's 1,024 rows).
int64_t adi_npad(int64_t n);
// `steps` Peaceman-Rachford steps (the grid in U is current when they end).
hipError_t launch_adi_steps(const AdiArgs& g, int steps);

// Relative residual per system (kern_util.hip).
hipError_t launch_residual(const double* a, const double* b, const double* c, const double* d,
                           const double* x, const int64_t* offsets, int64_t N, double* out,
                           hipStream_t stream);

}  // namespace trid
