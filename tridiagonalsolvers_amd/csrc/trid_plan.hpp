// trid_plan.hpp -- request validation and kernel planning (host only).
//
// A plan names the kernel family and its launch shape for one request. It is
// a pure function of (method, systemSize, numSystems, element size, options):
// it never looks at what else is in flight, so a system's result does not
// depend on the batch it arrives in.
#pragma once

#include <stdint.h>

#include "trid.h"

namespace trid {

enum class Family : int {
    Reg = 0,       // kern_reg.hip: register-resident CR / CR-PCR / PCR
    Lds = 1,       // kern_lds.hip: LDS-resident CR / CR-PCR with a pad map
    PcrThomas = 2, // kern_pcrgauss.hip: PCR levels then Thomas per subsystem
    Part = 3,      // kern_part.hip: CR / CR-PCR by separators, chunks of 1,024
};

struct Plan {
    Family family;
    int64_t n_pad;  // power of two >= n
    int M;          // Reg: unknowns per lane
    int L;          // Reg: lanes per system
    int rest;       // Reg / Lds: PCR system size (2 = plain CR)
    int pad;        // Lds: pad period in elements (0 = none)
    int gaussian;   // PcrThomas: number of Thomas subsystems
    int dense;      // 1 if n == n_pad (vector path possible)
    int io;         // Reg: dense-path I/O mode (RegArgs::io)
    int two_pass;   // PcrThomas: global PCR pass + strided Thomas pass
};

// Returns TRID_OK and fills *p, or TRID_EINVAL / TRID_ENOTSUP.
int make_plan(int method, int64_t n, int64_t N, int elem_size, const trid_opts& o, Plan* p);

// Plan of the FP32-solve / FP64-refinement kernels (register family only:
// CR, CR-PCR, PCR up to n = 1024, CR / CR-PCR up to n = 8192).
int make_refine_plan(int method, int64_t n, int64_t N, const trid_opts& o, Plan* p);

}  // namespace trid
