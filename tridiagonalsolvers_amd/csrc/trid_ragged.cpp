// trid_ragged.cpp -- ragged-batch plans: size buckets, the work queue of the
// persistent kernel (kern_ragged.hip) and its hipGraph.
//
// No counterpart in the reference (one systemSize per call). The plan is the
// format boundary between callers (the request server, users) and the
// kernel: host offsets in, a device work queue out. It depends only on the
// sizes and the bucket table, never on the data.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "trid.h"
#include "trid_launch.hpp"

struct trid_ragged_plan_s {
    int64_t N = 0;
    int64_t total = 0;
    int device = 0;
    int grid = 0;
    int n_wide_items = 0;
    int n_full_items = 0;
    int n_wg_items = 0;
    int n_wave_items = 0;
    int n_buckets = 0;
    trid::RaggedGrids grids{};  // of the refine mode, the FP64 mode and the
    trid::RaggedGrids grids_direct{};  // refine mode with the residual check
    trid::RaggedGrids grids_tol{};
    std::vector<trid::RaggedBucket> buckets;  // host copy (info)
    std::vector<int64_t> bucket_systems;
    char* dev = nullptr;  // one device block: systems, items, buckets, counters
    trid::RaggedSys* d_sys = nullptr;
    trid::RaggedItem* d_items = nullptr;
    trid::RaggedBucket* d_buckets = nullptr;
    int* d_counters = nullptr;
    // captured graph and what it was captured for
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void* key[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int key_iters = -1;  // iters + 16 * (FP64 mode)
    double key_tol = 0.0;
};

namespace {

bool is_pow2(int64_t v) { return v > 0 && (v & (v - 1)) == 0; }

int log2i(int64_t v)
{
    int k = 0;
    while ((int64_t(1) << k) < v) ++k;
    return k;
}

}  // namespace

namespace trid {

// Default configuration of a bucket of padded size n_pad.
RaggedBucket ragged_default_bucket(int n_pad)
{
    RaggedBucket b{};
    b.n_pad = n_pad;
    if (n_pad > 8192) {
        b.per_lane = 16;
        b.lanes = n_pad / 16;
        b.rest = b.lanes / 64;
    } else if (n_pad == 8192) {
        b.per_lane = 16;
        b.lanes = 512;
        b.rest = 512;  // one PCR equation per lane of the workgroup
    } else if (n_pad <= 512) {
        b.per_lane = n_pad / 2 < 8 ? n_pad / 2 : 8;
        b.lanes = n_pad / b.per_lane;
        b.rest = b.lanes < n_pad / 2 ? b.lanes : n_pad / 2;
        if (b.rest < 2) b.rest = 2;
    } else {
        b.per_lane = 8;
        b.lanes = n_pad / 8;
        b.rest = b.lanes / 64;  // one PCR equation per wave
        if (b.rest < 2) b.rest = 2;
    }
    return b;
}

int ragged_check_bucket(const RaggedBucket& b)
{
    if (!is_pow2(b.n_pad) || b.n_pad < 2 || b.n_pad > kRaggedMaxSize) return TRID_EINVAL;
    if (!is_pow2(b.per_lane) || !is_pow2(b.lanes) || !is_pow2(b.rest)) return TRID_EINVAL;
    if ((int64_t)b.per_lane * b.lanes != b.n_pad) return TRID_EINVAL;
    if (b.rest < 2 || b.rest > b.n_pad / 2 + (b.n_pad == 2 ? 1 : 0)) return TRID_EINVAL;
    if (b.n_pad <= 512) {
        // wave items: L <= 64 lanes, M <= 8, PCR on at most one equation per lane
        if (b.lanes < 2 || b.lanes > 64 || b.per_lane > 8 || b.rest > b.lanes)
            return TRID_EINVAL;
    } else {
        // workgroup items of W = lanes/64 >= 2 waves: 16 unknowns per lane
        // from 8192 on (8 waves; 16 waves above 8192); 8 or 16 below (W <= 8)
        const bool m_ok = b.n_pad >= 8192 ? b.per_lane == 16
                                          : (b.per_lane == 8 || b.per_lane == 16);
        if (!m_ok || b.lanes < 128 || b.lanes > 1024 || b.rest > b.lanes) return TRID_EINVAL;
    }
    return TRID_OK;
}

int ragged_build_host(const int64_t* offsets, int64_t N, const trid_bucket_cfg* table,
                      int tableSize, RaggedHostPlan* out)
{
    if (!offsets || N < 1 || N > 0x7fffffff || tableSize < 0 || (tableSize > 0 && !table))
        return TRID_EINVAL;
    if (offsets[0] < 0) return TRID_EINVAL;
    std::vector<int> bucket_of(N);
    for (int64_t s = 0; s < N; ++s) {
        const int64_t n = offsets[s + 1] - offsets[s];
        if (n < 2) return TRID_EINVAL;
        if (n > kRaggedMaxSize) return TRID_ENOTSUP;
        bucket_of[s] = log2i(n);  // 1 .. 14
    }
    // bucket table: defaults, overridden by the caller's (autotuned) entries
    RaggedBucket cfg[15];
    for (int k = 1; k <= 14; ++k) cfg[k] = ragged_default_bucket(1 << k);
    for (int t = 0; t < tableSize; ++t) {
        RaggedBucket b{table[t].size, table[t].per_lane, table[t].lanes, table[t].rest};
        if (ragged_check_bucket(b) != TRID_OK) return TRID_EINVAL;
        cfg[log2i(b.n_pad)] = b;
    }
    // systems grouped by bucket, largest bucket first, stable by index
    out->order.resize(N);
    for (int64_t s = 0; s < N; ++s) out->order[s] = (int32_t)s;
    std::stable_sort(out->order.begin(), out->order.end(),
                     [&](int32_t x, int32_t y) { return bucket_of[x] > bucket_of[y]; });
    // workgroup items by kernel: wide (above 8192 unknowns), 16 unknowns per
    // lane (8-wave workgroups, one per CU), 8 per lane (8-wave workgroups, two
    // per CU); each list largest bucket first
    std::vector<RaggedItem> wide_items, full_items, half_items, wave_items;
    out->buckets.clear();
    out->bucket_systems.clear();
    int64_t pos = 0;
    while (pos < N) {
        const int k = bucket_of[out->order[pos]];
        int64_t end = pos;
        while (end < N && bucket_of[out->order[end]] == k) ++end;
        const RaggedBucket& b = cfg[k];
        const int bi = (int)out->buckets.size();
        out->buckets.push_back(b);
        out->bucket_systems.push_back(end - pos);
        // wave items: 64/L systems; workgroup items: one wide system, or
        // 8/W systems of W <= 8 waves (a 512-thread workgroup)
        const int W = b.lanes / 64;
        const int per_item = b.n_pad <= 512 ? 64 / b.lanes : (b.n_pad > 8192 ? 1 : 8 / W);
        std::vector<RaggedItem>& dst = b.n_pad <= 512  ? wave_items
                                       : b.n_pad > 8192 ? wide_items
                                       : b.per_lane == 16 ? full_items
                                                          : half_items;
        for (int64_t f = pos; f < end; f += per_item) {
            RaggedItem it{};
            it.bucket = bi;
            it.first = (int32_t)f;
            it.count = (int32_t)std::min<int64_t>(per_item, end - f);
            dst.push_back(it);
        }
        pos = end;
    }
    out->n_wide_items = (int)wide_items.size();
    out->n_full_items = out->n_wide_items + (int)full_items.size();
    out->n_wg_items = out->n_full_items + (int)half_items.size();
    out->n_wave_items = (int)wave_items.size();
    out->items = wide_items;
    out->items.insert(out->items.end(), full_items.begin(), full_items.end());
    out->items.insert(out->items.end(), half_items.begin(), half_items.end());
    out->items.insert(out->items.end(), wave_items.begin(), wave_items.end());
    if (out->buckets.size() > (size_t)kRaggedMaxBuckets) return TRID_EINVAL;
    out->sys.resize(N);
    for (int64_t k = 0; k < N; ++k) {
        const int32_t s = out->order[k];
        out->sys[k] = RaggedSys{offsets[s], offsets[s + 1] - offsets[s]};
    }
    return TRID_OK;
}

// ---------------------------------------------------------------------------
// The autotune table file: JSON written by tridiagonalsolvers_amd.autotune,
// an object whose "buckets" array holds one object per bucket with integer
// "size", "per_lane", "lanes" and "rest" (other keys are measurements and are
// skipped). A small recursive-descent reader of that JSON, so that C and C++
// hosts load the same table the Python binding does.
namespace {

struct JsonReader {
    const char* p;
    const char* end;
    bool ok = true;

    void ws()
    {
        while (p < end && std::isspace((unsigned char)*p)) ++p;
    }
    bool eat(char ch)
    {
        ws();
        if (p < end && *p == ch) {
            ++p;
            return true;
        }
        return false;
    }
    bool string(std::string* out)
    {
        ws();
        if (p >= end || *p != '"') return ok = false;
        ++p;
        std::string v;
        while (p < end && *p != '"') {
            if (*p == '\\') {
                ++p;
                if (p >= end) return ok = false;
                if (*p == 'u') {  // keys of interest are ASCII: keep a placeholder
                    if (end - p < 5) return ok = false;
                    p += 4;
                    v.push_back('?');
                } else {
                    v.push_back(*p);
                }
                ++p;
                continue;
            }
            v.push_back(*p++);
        }
        if (p >= end) return ok = false;
        ++p;
        if (out) *out = v;
        return true;
    }
    bool number(double* out)
    {
        ws();
        const char* s0 = p;
        if (p < end && (*p == '-' || *p == '+')) ++p;
        while (p < end && (std::isdigit((unsigned char)*p) || *p == '.' || *p == 'e' ||
                           *p == 'E' || *p == '-' || *p == '+'))
            ++p;
        if (p == s0) return ok = false;
        std::string tok(s0, p);
        char* e = nullptr;
        const double v = std::strtod(tok.c_str(), &e);
        if (!e || *e) return ok = false;
        if (out) *out = v;
        return true;
    }
    // skip any value (nesting bounded by the file)
    bool skip(int depth = 0)
    {
        if (depth > 64) return ok = false;
        ws();
        if (p >= end) return ok = false;
        if (*p == '"') return string(nullptr);
        if (*p == '{') {
            ++p;
            if (eat('}')) return true;
            do {
                if (!string(nullptr) || !eat(':') || !skip(depth + 1)) return ok = false;
            } while (eat(','));
            return eat('}') || (ok = false);
        }
        if (*p == '[') {
            ++p;
            if (eat(']')) return true;
            do {
                if (!skip(depth + 1)) return ok = false;
            } while (eat(','));
            return eat(']') || (ok = false);
        }
        for (const char* lit : {"true", "false", "null"}) {
            const size_t n = std::strlen(lit);
            if ((size_t)(end - p) >= n && std::strncmp(p, lit, n) == 0) {
                p += n;
                return true;
            }
        }
        return number(nullptr);
    }
};

bool as_int32(double v, int32_t* out)
{
    if (!(v >= -2147483648.0 && v <= 2147483647.0) || v != (double)(int64_t)v) return false;
    *out = (int32_t)v;
    return true;
}

// One bucket object: the four integers, all required.
bool read_bucket(JsonReader& r, trid_bucket_cfg* b)
{
    if (!r.eat('{')) return false;
    bool have[4] = {false, false, false, false};
    static const char* keys[4] = {"size", "per_lane", "lanes", "rest"};
    int32_t* dst[4] = {&b->size, &b->per_lane, &b->lanes, &b->rest};
    if (!r.eat('}')) {
        do {
            std::string key;
            if (!r.string(&key) || !r.eat(':')) return false;
            int k = -1;
            for (int i = 0; i < 4; ++i)
                if (key == keys[i]) k = i;
            if (k < 0) {
                if (!r.skip()) return false;
                continue;
            }
            double v = 0;
            if (!r.number(&v) || !as_int32(v, dst[k])) return false;
            have[k] = true;
        } while (r.eat(','));
        if (!r.eat('}')) return false;
    }
    return have[0] && have[1] && have[2] && have[3];
}

}  // namespace

}  // namespace trid

extern "C" {

int trid_bucket_table_load(const char* path, trid_bucket_cfg* table, int capacity, int* count)
{
    if (!path || !count || capacity < 0 || (capacity > 0 && !table)) return TRID_EINVAL;
    *count = 0;
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return TRID_EINVAL;
    std::string text;
    char buf[65536];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) {
        text.append(buf, got);
        if (text.size() > (size_t(64) << 20)) {  // a table is a few KB
            std::fclose(f);
            return TRID_EINVAL;
        }
    }
    std::fclose(f);
    trid::JsonReader r{text.data(), text.data() + text.size()};
    if (!r.eat('{')) return TRID_EINVAL;
    bool found = false;
    int n = 0;
    if (!r.eat('}')) {
        do {
            std::string key;
            if (!r.string(&key) || !r.eat(':')) return TRID_EINVAL;
            if (key != "buckets") {
                if (!r.skip()) return TRID_EINVAL;
                continue;
            }
            if (found || !r.eat('[')) return TRID_EINVAL;
            found = true;
            if (r.eat(']')) continue;
            do {
                trid_bucket_cfg b{};
                if (!trid::read_bucket(r, &b)) return TRID_EINVAL;
                const trid::RaggedBucket rb{b.size, b.per_lane, b.lanes, b.rest};
                if (trid::ragged_check_bucket(rb) != TRID_OK) return TRID_EINVAL;
                if (n < capacity) table[n] = b;
                ++n;
            } while (r.eat(','));
            if (!r.eat(']')) return TRID_EINVAL;
        } while (r.eat(','));
        if (!r.eat('}')) return TRID_EINVAL;
    }
    r.ws();
    if (!found || r.p != r.end) return TRID_EINVAL;
    *count = n;
    return TRID_OK;
}

int trid_ragged_plan_create(trid_ragged_plan* out, const int64_t* offsets, int64_t numSystems,
                            const trid_bucket_cfg* table, int tableSize)
{
    if (!out) return TRID_EINVAL;
    *out = nullptr;
    trid::RaggedHostPlan hp;
    int st = trid::ragged_build_host(offsets, numSystems, table, tableSize, &hp);
    if (st != TRID_OK) return st;
    trid_ragged_plan_s* p = new (std::nothrow) trid_ragged_plan_s();
    if (!p) return TRID_ENOMEM;
    p->N = numSystems;
    p->total = offsets[numSystems];
    p->buckets = hp.buckets;
    p->bucket_systems = hp.bucket_systems;
    p->n_buckets = (int)hp.buckets.size();
    p->n_wide_items = hp.n_wide_items;
    p->n_full_items = hp.n_full_items;
    p->n_wg_items = hp.n_wg_items;
    p->n_wave_items = hp.n_wave_items;
    const std::vector<trid::RaggedItem>& items = hp.items;
    const std::vector<trid::RaggedBucket>& used = hp.buckets;
    // device block
    auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t b_sys = up(sizeof(trid::RaggedSys) * numSystems);
    const size_t b_itm = up(sizeof(trid::RaggedItem) * (items.size() + 1));
    const size_t b_bkt = up(sizeof(trid::RaggedBucket) * (used.size() + 1));
    const size_t b_ctr = 256;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p->dev), b_sys + b_itm + b_bkt + b_ctr);
    if (e != hipSuccess) {
        delete p;
        return TRID_ENOMEM;
    }
    p->d_sys = reinterpret_cast<trid::RaggedSys*>(p->dev);
    p->d_items = reinterpret_cast<trid::RaggedItem*>(p->dev + b_sys);
    p->d_buckets = reinterpret_cast<trid::RaggedBucket*>(p->dev + b_sys + b_itm);
    p->d_counters = reinterpret_cast<int*>(p->dev + b_sys + b_itm + b_bkt);
    if ((e = hipMemcpy(p->d_sys, hp.sys.data(), sizeof(trid::RaggedSys) * numSystems,
                       hipMemcpyHostToDevice)) != hipSuccess ||
        (!items.empty() &&
         (e = hipMemcpy(p->d_items, items.data(), sizeof(trid::RaggedItem) * items.size(),
                        hipMemcpyHostToDevice)) != hipSuccess) ||
        (e = hipMemcpy(p->d_buckets, used.data(), sizeof(trid::RaggedBucket) * used.size(),
                       hipMemcpyHostToDevice)) != hipSuccess) {
        (void)hipFree(p->dev);
        delete p;
        return TRID_EHIP;
    }
    (void)hipGetDevice(&p->device);
    trid::RaggedArgs g{};
    g.n_wide_items = p->n_wide_items;
    g.n_full_items = p->n_full_items;
    g.n_wg_items = p->n_wg_items;
    g.n_wave_items = p->n_wave_items;
    p->grids = trid::ragged_grids(p->device, g);
    g.tol = 1.0;
    p->grids_tol = trid::ragged_grids(p->device, g);
    g.tol = 0.0;
    g.direct = 1;
    p->grids_direct = trid::ragged_grids(p->device, g);
    p->grid = std::max(std::max(p->grids.wg, p->grids.half),
                       std::max(p->grids.wide, p->grids.wave));
    *out = p;
    return TRID_OK;
}

int trid_ragged_plan_destroy(trid_ragged_plan plan)
{
    if (!plan) return TRID_EINVAL;
    if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
    if (plan->graph) (void)hipGraphDestroy(plan->graph);
    if (plan->dev) (void)hipFree(plan->dev);
    delete plan;
    return TRID_OK;
}

int trid_ragged_plan_info(trid_ragged_plan plan, trid_ragged_info* info, trid_bucket_cfg* buckets,
                          int64_t* bucketSystems, int capacity)
{
    if (!plan || !info) return TRID_EINVAL;
    info->num_systems = plan->N;
    info->total_unknowns = plan->total;
    info->num_buckets = plan->n_buckets;
    info->workgroup_items = plan->n_wg_items;
    info->wave_items = plan->n_wave_items;
    info->grid = plan->grid;
    for (int i = 0; i < plan->n_buckets && i < capacity; ++i) {
        if (buckets) {
            buckets[i].size = plan->buckets[i].n_pad;
            buckets[i].per_lane = plan->buckets[i].per_lane;
            buckets[i].lanes = plan->buckets[i].lanes;
            buckets[i].rest = plan->buckets[i].rest;
        }
        if (bucketSystems) bucketSystems[i] = plan->bucket_systems[i];
    }
    return TRID_OK;
}

namespace {

int ragged_solve(trid_ragged_plan plan, const double* a, const double* b, const double* c,
                 const double* d, double* x, int iters, double tol, int flags, void* stream)
{
    if (!plan || !a || !b || !c || !d || !x || iters < 0 || iters > 8) return TRID_EINVAL;
    if (flags & ~(TRID_RAGGED_NO_GRAPH | TRID_RAGGED_FP64)) return TRID_EINVAL;
    const int direct = (flags & TRID_RAGGED_FP64) ? 1 : 0;
    if (direct && iters != 0) return TRID_EINVAL;
    if (direct && plan->n_wide_items > 0) return TRID_ENOTSUP;
    trid::RaggedArgs g{};
    g.a = a;
    g.b = b;
    g.c = c;
    g.d = d;
    g.x = x;
    g.sys = plan->d_sys;
    g.items = plan->d_items;
    g.buckets = plan->d_buckets;
    g.counters = plan->d_counters;
    g.n_wide_items = plan->n_wide_items;
    g.n_full_items = plan->n_full_items;
    g.n_wg_items = plan->n_wg_items;
    g.n_wave_items = plan->n_wave_items;
    g.n_buckets = plan->n_buckets;
    g.iters = iters;
    g.direct = direct;
    g.tol = tol;
    const trid::RaggedGrids& grids =
        direct ? plan->grids_direct : (tol > 0.0 ? plan->grids_tol : plan->grids);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if ((flags & TRID_RAGGED_NO_GRAPH) || s == nullptr) {
        hipError_t e = trid::launch_ragged(g, grids, s);
        return e == hipSuccess ? TRID_OK : TRID_EHIP;
    }
    const void* key[5] = {a, b, c, d, x};
    const int key_mode = iters + 16 * direct;
    const bool same = plan->exec && plan->key_iters == key_mode && plan->key_tol == tol &&
                      std::memcmp(key, plan->key, sizeof(key)) == 0;
    if (!same) {
        if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
        if (plan->graph) (void)hipGraphDestroy(plan->graph);
        plan->exec = nullptr;
        plan->graph = nullptr;
        hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return TRID_EHIP;
        hipError_t le = trid::launch_ragged(g, grids, s);
        e = hipStreamEndCapture(s, &plan->graph);
        if (le != hipSuccess || e != hipSuccess) {
            if (plan->graph) (void)hipGraphDestroy(plan->graph);
            plan->graph = nullptr;
            return TRID_EHIP;
        }
        e = hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            plan->exec = nullptr;
            return TRID_EHIP;
        }
        std::memcpy(plan->key, key, sizeof(key));
        plan->key_iters = key_mode;
        plan->key_tol = tol;
    }
    return hipGraphLaunch(plan->exec, s) == hipSuccess ? TRID_OK : TRID_EHIP;
}

}  // namespace

int trid_ragged_solve_f64(trid_ragged_plan plan, const double* a, const double* b, const double* c,
                          const double* d, double* x, int iters, int flags, void* stream)
{
    return ragged_solve(plan, a, b, c, d, x, iters, 0.0, flags, stream);
}

int trid_ragged_solve_tol_f64(trid_ragged_plan plan, const double* a, const double* b,
                              const double* c, const double* d, double* x, int maxIters,
                              double tol, int flags, void* stream)
{
    // tol: finite and >= 0 (0: exactly maxIters refinements); refine mode only
    if (!(tol >= 0.0) || tol > 1e308 || (flags & TRID_RAGGED_FP64)) return TRID_EINVAL;
    return ragged_solve(plan, a, b, c, d, x, maxIters, tol, flags, stream);
}

}  // extern "C"
