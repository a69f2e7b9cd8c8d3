// trid_plan.cpp -- request validation and kernel planning (host only, no HIP
// calls, so the rules are testable without a GPU through trid_plan_query).
#include "trid_plan.hpp"

namespace trid {

namespace {

bool is_pow2(int64_t v) { return v > 0 && (v & (v - 1)) == 0; }

int64_t next_pow2(int64_t v)
{
    int64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Largest LDS footprint one workgroup may use on gfx950 (160 KiB).
constexpr int64_t kLdsBytes = 160 * 1024;

// Dense-path I/O mode of the register kernels when the caller does not pick
// one (RegArgs::io numbering).
constexpr int kDefaultIo = 3;  // wave-coalesced through LDS, non-temporal

// Register family: unknowns per lane M and lanes per system L with
// L * M = n_pad. L <= 64: one wave holds a system (several systems share a
// wave when L < 64). L = 64 * W, W = 2..16: a workgroup of W waves holds one
// system (M <= 16 for FP32, M <= 8 for FP64, to fit 1024 threads' registers).
// Small batches trade unknowns per lane for lanes while a system still fits
// one wave, until the launch has 1024 waves (4 per CU): 512 x 512 runs at
// M = 8 (6.6 us, against 7.3 us at M = 16).
int plan_reg(int64_t n_pad, int64_t N, int elem_size, const trid_opts& o, Plan* p)
{
    int M, L;
    const int mmax_mw = elem_size == 4 ? 16 : 8;
    if (o.per_lane != 0 || o.lanes != 0) {
        if (o.per_lane == 0 || o.lanes == 0) return TRID_EINVAL;
        M = o.per_lane;
        L = o.lanes;
        if (!is_pow2(M) || M > 16 || !is_pow2(L) || L < 2 || L > 1024) return TRID_EINVAL;
        if ((int64_t)M * L != n_pad) return TRID_EINVAL;
        if (L > 64 && (M < 2 || M > 16)) return TRID_EINVAL;
        // FP64 at 16 unknowns per lane: 512-thread workgroups (at most 8 waves)
        if (L > 64 && M > mmax_mw && L > 512) return TRID_EINVAL;
    } else if (n_pad <= 64 * 16) {
        // Up to 16 unknowns per lane, at least two lanes per system.
        int64_t m = n_pad / 2;
        if (m > 16) m = 16;
        if (n_pad / m > 64) m = n_pad / 64;
        while (m > 2 && n_pad / (m / 2) <= 64 && N * (n_pad / m) < 1024 * 64) m /= 2;
        M = (int)m;
        L = (int)(n_pad / m);
    } else {
        M = mmax_mw;
        // FP64 from 4096 unknowns: 16 per lane in 512-thread workgroups
        // (8192: 0.39 ms against 0.44 ms at 8 per lane over 16 waves; 4096:
        // 0.31 against 0.34 ms, 8,192 and 4,096 systems)
        if (elem_size == 8 && n_pad >= 4096 && n_pad / 16 <= 512) M = 16;
        if (n_pad / M > 1024) return TRID_ENOTSUP;
        L = (int)(n_pad / M);
    }
    if (o.io_mode < 0 || o.io_mode > 4) return TRID_EINVAL;
    p->family = Family::Reg;
    p->M = M;
    p->L = L;
    p->io = o.io_mode > 0 ? o.io_mode - 1 : kDefaultIo;
    return TRID_OK;
}

// LDS-resident family: four arrays of n_pad elements (x overwrites d) plus
// one pad element per `pad` elements must fit in one workgroup's LDS.
// Separator form (kern_part.hip) for systems longer than the register
// kernels hold: chunks of 1,024 rows, one wave each (16 unknowns per lane),
// solved with switch size `rest` (at most half a chunk).
int plan_part(int rest, Plan* p)
{
    if (rest > 512) return TRID_ENOTSUP;
    p->family = Family::Part;
    p->M = 16;
    p->L = 64;
    p->rest = rest;
    p->io = kDefaultIo;
    return TRID_OK;
}

int plan_lds(int64_t n_pad, int elem_size, int pad, Plan* p)
{
    const int64_t padded = pad > 0 ? n_pad + n_pad / pad : n_pad;
    if (4 * padded * elem_size > kLdsBytes) return TRID_ENOTSUP;
    // PCR levels in LDS: at most 16 equations per thread of 256
    if (p->rest > 16 * 256) return TRID_ENOTSUP;
    p->family = Family::Lds;
    p->pad = pad;
    return TRID_OK;
}

}  // namespace

int make_plan(int method, int64_t n, int64_t N, int elem_size, const trid_opts& o, Plan* p)
{
    if (n < 2 || N < 0) return TRID_EINVAL;
    if (elem_size != 4 && elem_size != 8) return TRID_EINVAL;
    if (n > (int64_t(1) << 30)) return TRID_ENOTSUP;
    *p = Plan{};
    p->n_pad = next_pow2(n);
    p->dense = (p->n_pad == n) ? 1 : 0;
    const int64_t n_pad = p->n_pad;
    if (method == TRID_AUTO) method = TRID_CRPCR;

    switch (method) {
        case TRID_CR:
        case TRID_PCR:
        case TRID_CRPCR: {
            int rest;
            if (method == TRID_CR) {
                rest = 2;
            } else if (method == TRID_PCR) {
                rest = (int)(n_pad > (1 << 30) ? (1 << 30) : n_pad);
            } else if (o.rest == 0) {
                rest = -1;  // chosen below
            } else {
                if (!is_pow2(o.rest) || o.rest < 2 || o.rest > n_pad / 2) return TRID_EINVAL;
                rest = o.rest;
            }
            int st = plan_reg(n_pad, N, elem_size, o, p);
            if (st == TRID_EINVAL) return st;
            if (st == TRID_ENOTSUP) {
                // longer than a workgroup's registers hold: CR / CR-PCR by
                // separators (PCR has no such form)
                if (method == TRID_PCR) return TRID_ENOTSUP;
                return plan_part(rest < 0 ? 64 : rest, p);
            }
            if (st == TRID_OK && method == TRID_PCR && o.lanes == 0 && p->L < 64 &&
                n_pad >= 128) {
                // PCR: one system per wave, whose one-lane steps are DPP
                // shifts (131,072 x 512 FP32: 0.34 ms at 64 lanes x 8,
                // 0.38 ms at 32 x 16)
                p->L = 64;
                p->M = (int)(n_pad / 64);
            }
            if (st == TRID_OK) {
                if (rest < 0) {
                    // default switch size: one PCR equation per lane of the
                    // system's lanes (a wave, or the workgroup of the
                    // multi-wave kernels, whose PCR levels then take one
                    // barrier each and no cross-lane CR level; 4,096 x 8,192:
                    // FP64 339 us at rest = 512 against 354 us at 64, FP32
                    // 124 against 126 us)
                    const int64_t r = p->L;
                    rest = (int)(r < n_pad / 2 ? r : n_pad / 2);
                    if (rest < 2) rest = 2;
                }
                // multi-wave kernels hand PCR at most one equation per lane
                if (p->L <= 64 || rest <= p->L) {
                    p->rest = rest;
                    return TRID_OK;
                }
                if (o.lanes != 0) return TRID_EINVAL;
                *p = Plan{};
                p->n_pad = n_pad;
                p->dense = (n_pad == n) ? 1 : 0;
            }
            // Too long for one wave: LDS-resident kernel, unpadded map.
            if (rest < 0) rest = (int)(n_pad / 2 < 64 ? n_pad / 2 : 64);
            p->rest = rest;
            const int pad = o.lds_pad > 0 ? o.lds_pad : 0;
            return plan_lds(n_pad, elem_size, pad, p);
        }
        case TRID_CRNBC:
        case TRID_CRNBC32:
        case TRID_CRPCRNBC: {
            int rest = 2;
            if (method == TRID_CRPCRNBC) {
                if (o.rest == 0) {
                    rest = (int)(n_pad / 2 < 64 ? n_pad / 2 : 64);
                } else {
                    if (!is_pow2(o.rest) || o.rest < 2 || o.rest > n_pad / 2) return TRID_EINVAL;
                    rest = o.rest;
                }
            }
            int pad;
            if (method == TRID_CRNBC32) {
                pad = 0;  // the unpadded control (crNBC32banks.cu's map is the identity)
            } else if (o.lds_pad == 0) {
                // one pad element per 32: the fewest bank-conflict cycles of the
                // swept pads (8..128) in both precisions on gfx950
                pad = 32;
            } else if (o.lds_pad < 0) {
                pad = 0;
            } else {
                if (!is_pow2(o.lds_pad) || o.lds_pad < 2) return TRID_EINVAL;
                pad = o.lds_pad;
            }
            p->rest = rest;
            return plan_lds(n_pad, elem_size, pad, p);
        }
        case TRID_PCRGAUSS: {
            int G = o.gaussian;
            if (G == 0) G = (int)(n_pad / 2 < 64 ? n_pad / 2 : 64);
            if (!is_pow2(G) || G > n_pad / 2) return TRID_EINVAL;
            p->gaussian = G;
            p->family = Family::PcrThomas;
            // in-LDS form: four padded arrays in 160 KiB, <= 8 rows per thread
            const int pad = 64 * 4 / elem_size;
            const int64_t padded = n_pad + n_pad / pad;
            const bool fits = 4 * padded * elem_size <= kLdsBytes && n_pad <= 8 * 1024;
            p->two_pass = (!fits || (o.flags & TRID_FLAG_TWO_PASS)) ? 1 : 0;
            p->pad = p->two_pass ? 0 : pad;
            if (p->two_pass && G > (elem_size == 4 ? 1024 : 512)) return TRID_ENOTSUP;
            return TRID_OK;
        }
        default:
            return TRID_EINVAL;
    }
}

int make_refine_plan(int method, int64_t n, int64_t N, const trid_opts& o, Plan* p)
{
    if (method == TRID_AUTO) method = TRID_CRPCR;
    if (method != TRID_CR && method != TRID_CRPCR && method != TRID_PCR) return TRID_ENOTSUP;
    trid_opts q = o;
    const int64_t n_pad = next_pow2(n < 2 ? 2 : n);
    if (n_pad > 8192) return TRID_ENOTSUP;
    if (q.lanes == 0 && q.per_lane == 0) {
        // FP64 unknowns live in registers beside the FP32 equations: at most
        // 8 unknowns per lane (one wave up to n = 512, W waves above)
        const int64_t m = n_pad / 2 < 8 ? n_pad / 2 : 8;
        q.per_lane = (int32_t)m;
        q.lanes = (int32_t)(n_pad / m);
    }
    int st = make_plan(method, n, N, 4, q, p);
    if (st != TRID_OK) return st;
    if (p->family != Family::Reg) return TRID_ENOTSUP;
    if (p->L > 64 && p->M > 8) return TRID_EINVAL;
    return TRID_OK;
}

}  // namespace trid

extern "C" int trid_plan_query(int method, int64_t systemSize, int64_t numSystems,
                               int elemSize, const trid_opts* opts, trid_plan_info* out)
{
    if (!out) return TRID_EINVAL;
    trid_opts o{};
    if (opts) o = *opts;
    if (o.flags & ~TRID_FLAG_TWO_PASS) return TRID_EINVAL;
    trid::Plan p;
    int st = trid::make_plan(method, systemSize, numSystems, elemSize, o, &p);
    if (st != TRID_OK) return st;
    out->family = (int32_t)p.family;
    out->per_lane = p.M;
    out->lanes = p.L;
    out->rest = p.rest;
    out->lds_pad = p.pad;
    out->gaussian = p.gaussian;
    out->padded_size = p.n_pad;
    return TRID_OK;
}
