// trid_server.cpp -- continuous-batching request server over the ragged solve.
//
// No counterpart in the reference (a single-threaded test driver). Clients
// submit solve requests (one or more FP64 systems of any sizes 2..16384,
// host arrays) from any thread; one scheduler thread
//   * admits queued requests by device-memory footprint: a batch is formed
//     FIFO from what has arrived (requests that do not fit are passed over,
//     but a request that has waited longer than the batching window blocks
//     later ones, so nothing starves) and must fit the server's HBM budget
//     next to the resident operators and the batches still in flight;
//   * runs every batch as one ragged solve (FP32 CR-PCR + FP64 refinement,
//     kern_ragged.hip) on one of `slots` streams, so batches that arrive
//     while others run are admitted as soon as a slot and memory free up and
//     their copies and kernels overlap on the device;
//   * keeps registered operators (coefficient arrays) resident: requests on
//     an operator send only the right-hand sides, and the coefficients are
//     copied device-to-device into the batch;
//   * keeps ADI grids resident across time steps: an ADI request advances
//     its grid on a slot of its own (steps on one grid never overlap and keep
//     their submission order), interleaved with the solve batches; steps go
//     ahead of solve batches and skip the batching window, and while grids
//     exist (and there are two slots or more) slot 0 takes no solve batch, so
//     a step does not queue behind solves;
//   * reserves its whole budget once (one device allocation carved by a
//     first-fit allocator for batches and operators) and one pinned staging
//     buffer per slot, so serving allocates nothing per request;
//   * retires batches as their events complete and records per-request
//     timestamps (submit, launch, done) for latency percentiles.
// A system's result depends only on its own size (the plan is a pure function
// of the sizes), so a request gives the same answer alone and in any mix.
//
// The fake executor (TRID_SERVER_FAKE_EXECUTOR) makes no HIP call and
// computes nothing: batches "complete" after a modelled delay, optionally
// failing every k-th batch. It exists to test admission, batching,
// cancellation and failure handling without a device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "trid.h"
#include "trid_launch.hpp"
#include "trid_plan.hpp"

namespace {

using Clock = std::chrono::steady_clock;

double us_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::micro>(b - a).count();
}

enum ReqState { kQueued = 0, kRunning = 1, kDone = 2, kCancelled = 3, kFailed = 4 };

struct Req {
    uint64_t id = 0;
    trid_request r{};
    std::vector<int64_t> sizes;
    int64_t unknowns = 0;
    Clock::time_point t_submit, t_launch, t_done;
    int state = kQueued;
    int status = TRID_OK;
    int64_t batch = -1;
    int64_t batch_systems = 0;
    int64_t pos = 0;  // element offset inside its batch
    bool x_direct = false;  // result copied straight into a client host buffer
    int waiters = 0;        // threads inside trid_server_wait on this ticket
    bool retired = false;   // final state reported; no longer in the ticket map
};

struct Adi {
    int64_t n = 0;
    char* dev = nullptr;  // u | t | t2 (grids) | LU factors, inside the pool
    size_t pool_off = 0;
    size_t bytes = 0;
    int refs = 0;         // queued or running requests
    bool running = false;
};

struct Op {
    std::vector<int64_t> sizes;
    int64_t unknowns = 0;
    double* d_abc = nullptr;  // [a | b | c], unknowns each, inside the pool
    size_t pool_off = 0;
    size_t bytes = 0;
    int refs = 0;
};

// First-fit allocator over [0, capacity) (offsets only).
struct Pool {
    size_t capacity = 0;
    std::map<size_t, size_t> free_;  // offset -> size
    void init(size_t cap)
    {
        capacity = cap;
        free_.clear();
        if (cap) free_[0] = cap;
    }
    bool alloc(size_t bytes, size_t* off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes;
                const size_t start = it->first + bytes;
                free_.erase(it);
                if (rest) free_[start] = rest;
                return true;
            }
        }
        return false;
    }
    // From the top end of the highest free block that holds `bytes`:
    // resident objects (operators, ADI grids) pile up from the top of the
    // pool, so the space batches come and go in stays one block at the bottom.
    bool alloc_high(size_t bytes, size_t* off)
    {
        for (auto it = free_.rbegin(); it != free_.rend(); ++it) {
            if (it->second >= bytes) {
                const size_t start = it->first;
                const size_t size = it->second;
                *off = start + size - bytes;
                free_.erase(start);
                if (size > bytes) free_[start] = size - bytes;
                return true;
            }
        }
        return false;
    }
    size_t largest() const
    {
        size_t m = 0;
        for (const auto& kv : free_) m = std::max(m, kv.second);
        return m;
    }
    void release(size_t off, size_t bytes)
    {
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Host copies of a batch (into / out of pinned staging), split over up to 8
// threads: one thread moves ~10 GB/s, below what the PCIe link takes.
struct CopyJob {
    void* dst;
    const void* src;
    size_t bytes;
};

void parallel_copy(const std::vector<CopyJob>& jobs)
{
    size_t total = 0;
    for (const CopyJob& j : jobs) total += j.bytes;
    const size_t kChunk = size_t(4) << 20;
    int nt = (int)std::min<size_t>(8, total / kChunk);
    if (nt <= 1) {
        for (const CopyJob& j : jobs) std::memcpy(j.dst, j.src, j.bytes);
        return;
    }
    // cut every job into pieces, deal the pieces round-robin
    std::vector<CopyJob> pieces;
    for (const CopyJob& j : jobs)
        for (size_t o = 0; o < j.bytes; o += kChunk)
            pieces.push_back(CopyJob{static_cast<char*>(j.dst) + o,
                                     static_cast<const char*>(j.src) + o,
                                     std::min(kChunk, j.bytes - o)});
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&pieces, t, nt] {
            for (size_t i = t; i < pieces.size(); i += nt)
                std::memcpy(pieces[i].dst, pieces[i].src, pieces[i].bytes);
        });
    for (auto& x : th) x.join();
}

// Device (and staging) layout of one batch of N systems, `total` unknowns.
struct BatchLayout {
    size_t arr;   // bytes of one FP64 array
    size_t sys;   // systems in queue order
    size_t itm;   // items
    size_t bkt;   // buckets
    size_t ctr;   // counters
    size_t bytes;
    BatchLayout(int64_t total, int64_t N)
    {
        arr = align256((size_t)total * sizeof(double));
        sys = align256(sizeof(trid::RaggedSys) * N);
        itm = align256(sizeof(trid::RaggedItem) * (N + 1));  // at most one item per system
        bkt = align256(sizeof(trid::RaggedBucket) * 16);
        ctr = 256;
        bytes = 5 * arr + sys + itm + bkt + ctr;
    }
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    char* pinned = nullptr;
    size_t pinned_bytes = 0;
    bool busy = false;
    std::vector<Req*> reqs;
    int64_t batch = -1;
    int64_t total = 0, N = 0;
    size_t dev_off = 0, dev_bytes = 0;
    bool failed = false;
    Clock::time_point fake_done;
};

}  // namespace

struct trid_server_s {
    trid_server_opts o{};
    bool fake = false;
    int device = 0;
    std::mutex mu;
    std::condition_variable cv_sched;  // wakes the scheduler
    std::condition_variable cv_done;   // wakes waiters
    std::thread th;
    bool stopping = false;
    uint64_t next_id = 1;
    int next_op = 0;
    int64_t next_batch = 0;
    std::deque<Req*> pending;
    std::unordered_map<uint64_t, Req*> reqs;
    std::unordered_map<int, Op> ops;
    std::unordered_map<int, Adi> adis;
    int next_adi = 0;
    std::vector<Slot> slots;
    Pool pool;
    char* d_pool = nullptr;
    // page-locked client buffers handed out by trid_server_host_alloc:
    // request arrays inside them are copied by DMA with no staging copy
    std::map<const char*, size_t> host_bufs;
    bool in_host_buf(const void* p, size_t bytes) const
    {
        const char* c = static_cast<const char*>(p);
        auto it = host_bufs.upper_bound(c);
        if (it == host_bufs.begin()) return false;
        --it;
        return c >= it->first && c + bytes <= it->first + it->second;
    }
    size_t ops_bytes = 0;
    size_t in_use = 0;
    trid_server_stats st{};
    std::vector<trid_bucket_cfg> table;

    size_t budget_free() const
    {
        const size_t used = in_use + ops_bytes;
        return used >= (size_t)o.hbm_budget_bytes ? 0 : (size_t)o.hbm_budget_bytes - used;
    }
    void note_peak()
    {
        const int64_t used = (int64_t)(in_use + ops_bytes);
        if (used > st.peak_device_bytes) st.peak_device_bytes = used;
    }
    void finish(Req* q, int state, int status)
    {
        q->state = state;
        q->status = status;
        q->t_done = Clock::now();
        if (q->r.op >= 0) {
            auto it = ops.find(q->r.op);
            if (it != ops.end()) it->second.refs--;
        }
        if (q->r.adi >= 0) {
            auto it = adis.find(q->r.adi);
            if (it != adis.end()) {
                it->second.refs--;
                it->second.running = false;
            }
        }
        if (state == kDone) st.completed++;
        else if (state == kFailed) st.failed++;
        else if (state == kCancelled) st.cancelled++;
    }

    int launch(Slot& s, std::unique_lock<std::mutex>& lk);
    void retire(Slot& s, bool ok, std::unique_lock<std::mutex>& lk);
    void run();
};

// ---------------------------------------------------------------------------

// Called with the server mutex held (lk). A solve batch is planned under the
// lock (sizes, work queue, which arrays go by DMA, device addresses of the
// operators it uses); the lock is released for the host staging copies and
// the HIP calls, which touch only the slot (owned by the scheduler thread)
// and that snapshot, so submit / wait / stats from clients never wait for a
// batch's copies. The lock is held again on return.
int trid_server_s::launch(Slot& s, std::unique_lock<std::mutex>& lk)
{
    if (s.reqs.size() == 1 && s.reqs[0]->r.adi >= 0) {
        Req* q = s.reqs[0];
        Adi& a = adis[q->r.adi];
        a.running = true;
        if (fake) {
            s.fake_done = Clock::now() + std::chrono::microseconds(
                o.fake_batch_us > 0 ? (int64_t)o.fake_batch_us
                                    : 20 + (int64_t)(a.n * a.n * q->r.adi_steps / 50000));
            return TRID_OK;
        }
        trid::AdiArgs g{};
        const size_t grid = (size_t)a.n * a.n * sizeof(double);
        g.u = reinterpret_cast<double*>(a.dev);
        g.t = reinterpret_cast<double*>(a.dev + grid);
        g.t2 = reinterpret_cast<double*>(a.dev + 2 * grid);
        g.fac = reinterpret_cast<double*>(a.dev + align256(3 * grid));
        g.n = a.n;
        g.r = q->r.adi_r;
        g.stream = s.stream;
        if (trid::launch_adi_steps(g, q->r.adi_steps) != hipSuccess) return TRID_EHIP;
        return hipEventRecord(s.done, s.stream) == hipSuccess ? TRID_OK : TRID_EHIP;
    }
    const int64_t total = s.total, N = s.N;
    BatchLayout L(total, N);
    // concatenated sizes and plan
    std::vector<int64_t> offs(N + 1);
    offs[0] = 0;
    int64_t k = 0;
    for (Req* q : s.reqs) {
        q->pos = offs[k];
        for (int64_t n : q->sizes) {
            offs[k + 1] = offs[k] + n;
            ++k;
        }
    }
    trid::RaggedHostPlan hp;
    int st = trid::ragged_build_host(offs.data(), N, table.data(), (int)table.size(), &hp);
    if (st != TRID_OK) return st;
    if (fake) {
        const double us = o.fake_batch_us > 0 ? (double)o.fake_batch_us
                                              : 20.0 + (double)total / 5.0e4;  // ~50 unknowns/ns
        s.fake_done = Clock::now() + std::chrono::microseconds((int64_t)us);
        if (o.fake_fail_every > 0 && (s.batch % o.fake_fail_every) == o.fake_fail_every - 1)
            s.failed = true;
        return TRID_OK;
    }
    // stage inputs and descriptors in pinned memory, same layout as device
    char* P = s.pinned;
    double* pa = reinterpret_cast<double*>(P);
    double* pb = reinterpret_cast<double*>(P + L.arr);
    double* pc = reinterpret_cast<double*>(P + 2 * L.arr);
    double* pd = reinterpret_cast<double*>(P + 3 * L.arr);
    char* pdesc = P + 5 * L.arr;
    // arrays in client page-locked buffers go to the device by DMA straight
    // from there; the others are copied into the slot's staging block first
    struct Dma {
        int arr;
        int64_t pos;
        const void* src;
        size_t bytes;
    };
    struct OpCopy {
        int64_t pos;
        const double* abc;  // [a | b | c] of the operator, `stride` apart
        int64_t stride;
        size_t bytes;
    };
    struct XCopy {
        void* dst;
        int64_t pos;
        size_t bytes;
    };
    std::vector<CopyJob> jobs;
    std::vector<Dma> dmas;
    std::vector<OpCopy> opcopies;
    std::vector<XCopy> xcopies;
    for (Req* q : s.reqs) {
        const size_t nb = (size_t)q->unknowns * sizeof(double);
        const double* src[4] = {q->r.a, q->r.b, q->r.c, q->r.d};
        double* stg[4] = {pa, pb, pc, pd};
        for (int a = q->r.op < 0 ? 0 : 3; a < 4; ++a) {
            if (in_host_buf(src[a], nb)) {
                dmas.push_back(Dma{a, q->pos, src[a], nb});
            } else {
                jobs.push_back(CopyJob{stg[a] + q->pos, src[a], nb});
                dmas.push_back(Dma{a, q->pos, stg[a] + q->pos, nb});
            }
        }
        if (q->r.op >= 0) {
            const Op& op = ops[q->r.op];  // held by the request's reference
            opcopies.push_back(OpCopy{q->pos, op.d_abc, op.unknowns, nb});
        }
        // results: straight into client page-locked buffers, else via staging
        q->x_direct = in_host_buf(q->r.x, nb);
        xcopies.push_back(XCopy{q->x_direct ? static_cast<void*>(q->r.x)
                                            : static_cast<void*>(P + 4 * L.arr +
                                                                 q->pos * sizeof(double)),
                                q->pos, nb});
    }
    char* D = d_pool + s.dev_off;

    lk.unlock();
    parallel_copy(jobs);
    std::memcpy(pdesc, hp.sys.data(), sizeof(trid::RaggedSys) * N);
    std::memcpy(pdesc + L.sys, hp.items.data(), sizeof(trid::RaggedItem) * hp.items.size());
    std::memcpy(pdesc + L.sys + L.itm, hp.buckets.data(),
                sizeof(trid::RaggedBucket) * hp.buckets.size());
    hipError_t e = hipSuccess;
    for (const Dma& m : dmas) {
        if (e != hipSuccess) break;
        e = hipMemcpyAsync(D + m.arr * L.arr + m.pos * sizeof(double), m.src, m.bytes,
                           hipMemcpyHostToDevice, s.stream);
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(D + 5 * L.arr, pdesc, L.sys + L.itm + L.bkt,
                           hipMemcpyHostToDevice, s.stream);
    for (const OpCopy& c : opcopies)
        for (int a = 0; a < 3 && e == hipSuccess; ++a)
            e = hipMemcpyAsync(D + a * L.arr + c.pos * sizeof(double), c.abc + a * c.stride,
                               c.bytes, hipMemcpyDeviceToDevice, s.stream);
    if (e == hipSuccess) {
        trid::RaggedArgs g{};
        g.a = reinterpret_cast<const double*>(D);
        g.b = reinterpret_cast<const double*>(D + L.arr);
        g.c = reinterpret_cast<const double*>(D + 2 * L.arr);
        g.d = reinterpret_cast<const double*>(D + 3 * L.arr);
        g.x = reinterpret_cast<double*>(D + 4 * L.arr);
        char* desc = D + 5 * L.arr;
        g.sys = reinterpret_cast<const trid::RaggedSys*>(desc);
        g.items = reinterpret_cast<const trid::RaggedItem*>(desc + L.sys);
        g.buckets = reinterpret_cast<const trid::RaggedBucket*>(desc + L.sys + L.itm);
        g.counters = reinterpret_cast<int*>(desc + L.sys + L.itm + L.bkt);
        g.n_wide_items = hp.n_wide_items;
        g.n_full_items = hp.n_full_items;
        g.n_wg_items = hp.n_wg_items;
        g.n_wave_items = hp.n_wave_items;
        g.n_buckets = (int)hp.buckets.size();
        g.iters = o.iters;
        e = trid::launch_ragged(g, trid::ragged_grids(device, g), s.stream);
    }
    for (const XCopy& c : xcopies) {
        if (e != hipSuccess) break;
        e = hipMemcpyAsync(c.dst, D + 4 * L.arr + c.pos * sizeof(double), c.bytes,
                           hipMemcpyDeviceToHost, s.stream);
    }
    if (e == hipSuccess) e = hipEventRecord(s.done, s.stream);
    lk.lock();
    return e == hipSuccess ? TRID_OK : TRID_EHIP;
}

// Called with the server mutex held (lk); released while the results are
// copied out of the slot's staging block (the slot's requests are running,
// so no client call changes them meanwhile).
void trid_server_s::retire(Slot& s, bool ok, std::unique_lock<std::mutex>& lk)
{
    const bool adi_req = s.reqs.size() == 1 && s.reqs[0]->r.adi >= 0;
    if (ok && !fake && !adi_req) {
        BatchLayout L(s.total, s.N);
        const double* px = reinterpret_cast<const double*>(s.pinned + 4 * L.arr);
        std::vector<CopyJob> jobs;
        for (Req* q : s.reqs)
            if (!q->x_direct)
                jobs.push_back(
                    CopyJob{q->r.x, px + q->pos, (size_t)q->unknowns * sizeof(double)});
        lk.unlock();
        parallel_copy(jobs);
        lk.lock();
    }
    for (Req* q : s.reqs) finish(q, ok ? kDone : kFailed, ok ? TRID_OK : TRID_EHIP);
    if (s.dev_bytes) pool.release(s.dev_off, s.dev_bytes);
    in_use -= s.dev_bytes;
    s.reqs.clear();
    s.busy = false;
    s.failed = false;
    cv_done.notify_all();
}

void trid_server_s::run()
{
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        // ---- retire finished batches ------------------------------------
        bool any_busy = false;
        for (Slot& s : slots) {
            if (!s.busy) continue;
            if (fake) {
                if (Clock::now() >= s.fake_done) retire(s, !s.failed, lk);
            } else {
                const hipError_t q = hipEventQuery(s.done);
                if (q == hipSuccess) retire(s, true, lk);
                else if (q != hipErrorNotReady) retire(s, false, lk);
            }
            any_busy = any_busy || s.busy;
        }
        if (stopping && pending.empty() && !any_busy) break;

        // ---- form and launch batches on free slots ------------------------
        for (size_t si = 0; si < slots.size(); ++si) {
            Slot& s = slots[si];
            if (s.busy || pending.empty()) continue;
            const Clock::time_point now = Clock::now();
            // ADI steps first: each runs alone, so it does not wait for a
            // batching window; the first pending step of a grid that is not
            // running goes (one at a time per grid, in submission order)
            Req* adi_q = nullptr;
            {
                std::vector<int> seen;
                for (auto it = pending.begin(); it != pending.end(); ++it) {
                    const int g = (*it)->r.adi;
                    if (g < 0 || std::find(seen.begin(), seen.end(), g) != seen.end()) continue;
                    seen.push_back(g);
                    if (!adis[g].running) {
                        adi_q = *it;
                        pending.erase(it);
                        break;
                    }
                }
            }
            if (adi_q) {
                s.busy = true;
                s.reqs.assign(1, adi_q);
                s.total = 0;
                s.N = 0;
                s.dev_off = 0;
                s.dev_bytes = 0;
                s.batch = next_batch++;
                st.batches++;
                adi_q->state = kRunning;
                adi_q->t_launch = now;
                adi_q->batch = s.batch;
                adi_q->batch_systems = 2 * adis[adi_q->r.adi].n;
                const int rc = launch(s, lk);
                if (rc != TRID_OK) {
                    finish(adi_q, kFailed, rc);
                    s.reqs.clear();
                    s.busy = false;
                    cv_done.notify_all();
                }
                continue;
            }
            // with resident ADI grids and more than one slot, slot 0 is kept
            // for their steps, so a step never queues behind solve batches
            if (si == 0 && slots.size() >= 2 && !adis.empty()) continue;
            bool any_solve = false;
            for (Req* q : pending) any_solve = any_solve || q->r.adi < 0;
            if (!any_solve) continue;
            const bool device_busy = std::any_of(slots.begin(), slots.end(),
                                                 [](const Slot& x) { return x.busy; });
            // while the device is busy, let arrivals accumulate for up to the
            // batching window (unless a full batch is already waiting)
            int64_t waiting = 0;
            for (Req* q : pending) waiting += q->unknowns;
            const double age = us_between(pending.front()->t_submit, now);
            if (device_busy && age < o.batch_window_us && waiting < o.max_batch_unknowns &&
                !stopping)
                continue;
            // FIFO selection of solve requests within the batch limits and
            // the largest free block of the pool (a batch is one block)
            int64_t total = 0, N = 0;
            std::vector<Req*> pick;
            const size_t free_b = std::min(budget_free(), pool.largest());
            Req* head = nullptr;  // the oldest solve request
            for (auto it = pending.begin(); it != pending.end();) {
                Req* q = *it;
                if (q->r.adi >= 0) {  // waits for its grid (taken above when free)
                    ++it;
                    continue;
                }
                if (!head) head = q;
                const int64_t t2 = total + q->unknowns;
                const int64_t n2 = N + (int64_t)q->sizes.size();
                const bool fits = t2 <= o.max_batch_unknowns && n2 <= o.max_batch_systems &&
                                  BatchLayout(t2, n2).bytes <= free_b;
                if (fits) {
                    pick.push_back(q);
                    total = t2;
                    N = n2;
                    it = pending.erase(it);
                    continue;
                }
                // an old request that does not fit blocks the ones behind it
                if (us_between(q->t_submit, now) > o.batch_window_us) break;
                ++it;
            }
            if (pick.empty()) {
                // Nothing fits. With no solve batch in flight no retirement
                // will free memory (the free blocks are the gaps between
                // resident operators and grids), so the oldest request fails
                // with TRID_ENOMEM instead of waiting for ever.
                const bool batch_in_flight = std::any_of(
                    slots.begin(), slots.end(),
                    [](const Slot& x) { return x.busy && x.dev_bytes > 0; });
                if (!batch_in_flight && head) {
                    pending.erase(std::find(pending.begin(), pending.end(), head));
                    finish(head, kFailed, TRID_ENOMEM);
                    cv_done.notify_all();
                    continue;
                }
                break;
            }
            const BatchLayout L(total, N);
            size_t off = 0;
            if (!pool.alloc(L.bytes, &off)) {
                // fragmented: put them back in order and wait for retirements
                for (auto r = pick.rbegin(); r != pick.rend(); ++r) pending.push_front(*r);
                std::stable_sort(pending.begin(), pending.end(),
                                 [](Req* x, Req* y) { return x->id < y->id; });
                break;
            }
            s.busy = true;
            s.reqs = pick;
            s.total = total;
            s.N = N;
            s.dev_off = off;
            s.dev_bytes = L.bytes;
            s.batch = next_batch++;
            in_use += L.bytes;
            note_peak();
            st.batches++;
            for (Req* q : pick) {
                q->state = kRunning;
                q->t_launch = now;
                q->batch = s.batch;
                q->batch_systems = N;
            }
            const int rc = launch(s, lk);
            if (rc != TRID_OK) {
                for (Req* q : s.reqs) finish(q, kFailed, rc);
                pool.release(s.dev_off, s.dev_bytes);
                in_use -= s.dev_bytes;
                s.reqs.clear();
                s.busy = false;
                cv_done.notify_all();
            }
        }
        const bool busy_now = std::any_of(slots.begin(), slots.end(),
                                          [](const Slot& x) { return x.busy; });
        if (busy_now || !pending.empty())
            cv_sched.wait_for(lk, std::chrono::microseconds(20));
        else
            cv_sched.wait(lk, [&] { return stopping || !pending.empty(); });
    }
}

// ---------------------------------------------------------------------------

extern "C" {

int trid_server_create(trid_server* out, const trid_server_opts* opts)
{
    if (!out) return TRID_EINVAL;
    *out = nullptr;
    trid_server_opts o{};
    if (opts) o = *opts;
    if (o.hbm_budget_bytes == 0) o.hbm_budget_bytes = int64_t(4) << 30;
    if (o.slots == 0) o.slots = 3;
    if (o.max_batch_systems == 0) o.max_batch_systems = 65536;
    if (o.max_batch_unknowns == 0) o.max_batch_unknowns = int64_t(1) << 23;
    if (o.batch_window_us == 0) o.batch_window_us = 200;
    if (o.iters < 0) o.iters = 2;
    if (o.hbm_budget_bytes < 0 || o.slots < 1 || o.slots > 16 || o.max_batch_systems < 1 ||
        o.max_batch_unknowns < 2 || o.batch_window_us < 0 || o.iters > 8 ||
        (o.flags & ~TRID_SERVER_FAKE_EXECUTOR) || o.fake_fail_every < 0 || o.fake_batch_us < 0 ||
        (o.bucket_table_size > 0 && !o.bucket_table) || o.bucket_table_size < 0)
        return TRID_EINVAL;
    trid_server_s* s = new (std::nothrow) trid_server_s();
    if (!s) return TRID_ENOMEM;
    s->o = o;
    s->fake = (o.flags & TRID_SERVER_FAKE_EXECUTOR) != 0;
    for (int i = 0; i < o.bucket_table_size; ++i) s->table.push_back(o.bucket_table[i]);
    s->slots.resize(o.slots);
    const BatchLayout maxL(o.max_batch_unknowns, o.max_batch_systems);
    s->pool.init((size_t)o.hbm_budget_bytes);
    if (!s->fake) {
        int st = TRID_OK;
        if (hipGetDevice(&s->device) != hipSuccess) st = TRID_ENODEV;
        if (st == TRID_OK && hipMalloc(reinterpret_cast<void**>(&s->d_pool),
                                       (size_t)o.hbm_budget_bytes) != hipSuccess)
            st = TRID_ENOMEM;
        for (Slot& sl : s->slots) {
            if (st != TRID_OK) break;
            if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
                hipHostMalloc(reinterpret_cast<void**>(&sl.pinned), maxL.bytes,
                              hipHostMallocDefault) != hipSuccess)
                st = TRID_ENOMEM;
            sl.pinned_bytes = maxL.bytes;
        }
        if (st != TRID_OK) {
            trid_server_destroy(s);
            return st;
        }
    }
    s->st.budget_bytes = o.hbm_budget_bytes;
    s->th = std::thread([s] { s->run(); });
    *out = s;
    return TRID_OK;
}

int trid_server_destroy(trid_server s)
{
    if (!s) return TRID_EINVAL;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv_sched.notify_all();
    if (s->th.joinable()) s->th.join();
    for (Slot& sl : s->slots) {
        if (sl.pinned) (void)hipHostFree(sl.pinned);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    if (s->d_pool) (void)hipFree(s->d_pool);
    for (auto& kv : s->host_bufs) {
        if (s->fake) std::free(const_cast<char*>(kv.first));
        else (void)hipHostFree(const_cast<char*>(kv.first));
    }
    for (auto& kv : s->reqs) delete kv.second;
    delete s;
    return TRID_OK;
}

int trid_server_host_alloc(trid_server s, size_t bytes, void** ptr)
{
    if (!s || !ptr || bytes == 0) return TRID_EINVAL;
    *ptr = nullptr;
    void* p = nullptr;
    if (s->fake) {
        p = std::aligned_alloc(256, (bytes + 255) & ~(size_t)255);
        if (!p) return TRID_ENOMEM;
    } else if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
        return TRID_ENOMEM;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->host_bufs[static_cast<const char*>(p)] = bytes;
    *ptr = p;
    return TRID_OK;
}

int trid_server_host_free(trid_server s, void* ptr)
{
    if (!s || !ptr) return TRID_EINVAL;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->host_bufs.find(static_cast<const char*>(ptr));
        if (it == s->host_bufs.end()) return TRID_EINVAL;
        s->host_bufs.erase(it);
    }
    if (s->fake) std::free(ptr);
    else if (hipHostFree(ptr) != hipSuccess) return TRID_EHIP;
    return TRID_OK;
}

int trid_server_register_operator(trid_server s, const int64_t* sizes, int64_t numSystems,
                                  const double* a, const double* b, const double* c, int32_t* op)
{
    if (!s || !sizes || numSystems < 1 || !a || !b || !c || !op) return TRID_EINVAL;
    Op o;
    for (int64_t i = 0; i < numSystems; ++i) {
        if (sizes[i] < 2) return TRID_EINVAL;
        if (sizes[i] > trid::kRaggedMaxSize) return TRID_ENOTSUP;
        o.sizes.push_back(sizes[i]);
        o.unknowns += sizes[i];
    }
    o.bytes = align256(3 * (size_t)o.unknowns * sizeof(double));
    std::lock_guard<std::mutex> lk(s->mu);
    if (o.bytes > s->budget_free()) return TRID_ENOMEM;
    size_t off = 0;
    if (!s->pool.alloc_high(o.bytes, &off)) return TRID_ENOMEM;
    o.pool_off = off;
    if (!s->fake) {
        o.d_abc = reinterpret_cast<double*>(s->d_pool + off);
        const size_t nb = (size_t)o.unknowns * sizeof(double);
        if (hipMemcpy(o.d_abc, a, nb, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(o.d_abc + o.unknowns, b, nb, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(o.d_abc + 2 * o.unknowns, c, nb, hipMemcpyHostToDevice) != hipSuccess) {
            s->pool.release(off, o.bytes);
            return TRID_EHIP;
        }
    }
    s->ops_bytes += o.bytes;
    s->note_peak();
    *op = s->next_op++;
    s->ops[*op] = std::move(o);
    return TRID_OK;
}

int trid_server_release_operator(trid_server s, int32_t op)
{
    if (!s) return TRID_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->ops.find(op);
    if (it == s->ops.end()) return TRID_EINVAL;
    if (it->second.refs > 0) return TRID_EAGAIN;
    s->pool.release(it->second.pool_off, it->second.bytes);
    s->ops_bytes -= it->second.bytes;
    s->ops.erase(it);
    return TRID_OK;
}

int trid_server_create_adi(trid_server s, int64_t n, const double* u0, int32_t* id)
{
    if (!s || !u0 || !id || n < 2 || n > trid::kRaggedMaxSize) return TRID_EINVAL;
    Adi a;
    a.n = n;
    const size_t grid = (size_t)n * n * sizeof(double);
    a.bytes = align256(3 * grid) + align256((size_t)trid::adi_npad(n) * sizeof(double));
    std::lock_guard<std::mutex> lk(s->mu);
    if (a.bytes > s->budget_free()) return TRID_ENOMEM;
    size_t off = 0;
    if (!s->pool.alloc_high(a.bytes, &off)) return TRID_ENOMEM;
    a.pool_off = off;
    if (!s->fake) {
        a.dev = s->d_pool + off;
        if (hipMemcpy(a.dev, u0, grid, hipMemcpyHostToDevice) != hipSuccess) {
            s->pool.release(off, a.bytes);
            return TRID_EHIP;
        }
    }
    s->ops_bytes += a.bytes;
    s->note_peak();
    *id = s->next_adi++;
    s->adis[*id] = a;
    return TRID_OK;
}

int trid_server_read_adi(trid_server s, int32_t id, double* u)
{
    if (!s || !u) return TRID_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->adis.find(id);
    if (it == s->adis.end()) return TRID_EINVAL;
    if (it->second.refs > 0) return TRID_EAGAIN;
    if (s->fake) return TRID_OK;
    const size_t grid = (size_t)it->second.n * it->second.n * sizeof(double);
    return hipMemcpy(u, it->second.dev, grid, hipMemcpyDeviceToHost) == hipSuccess ? TRID_OK
                                                                                    : TRID_EHIP;
}

int trid_server_release_adi(trid_server s, int32_t id)
{
    if (!s) return TRID_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->adis.find(id);
    if (it == s->adis.end()) return TRID_EINVAL;
    if (it->second.refs > 0) return TRID_EAGAIN;
    s->pool.release(it->second.pool_off, it->second.bytes);
    s->ops_bytes -= it->second.bytes;
    s->adis.erase(it);
    return TRID_OK;
}

int trid_server_submit(trid_server s, const trid_request* req, uint64_t* ticket)
{
    if (!s || !req || !ticket) return TRID_EINVAL;
    if (req->kind < TRID_REQ_SOLVE || req->kind > TRID_REQ_ADI) return TRID_EINVAL;
    if (req->kind != TRID_REQ_ADI && (!req->d || !req->x)) return TRID_EINVAL;
    Req* q = new (std::nothrow) Req();
    if (!q) return TRID_ENOMEM;
    q->r = *req;
    // internal convention: op / adi are -1 unless the request is of that kind
    if (req->kind != TRID_REQ_OPERATOR) q->r.op = -1;
    if (req->kind != TRID_REQ_ADI) q->r.adi = -1;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->stopping) {
        delete q;
        return TRID_EINVAL;
    }
    if (req->kind == TRID_REQ_ADI) {
        auto it = s->adis.find(req->adi);
        if (it == s->adis.end() || req->adi_steps < 1 || !(req->adi_r > 0.0) ||
            req->adi_r > 1.0e6) {
            delete q;
            return TRID_EINVAL;
        }
        q->r.op = -1;
        it->second.refs++;
        q->id = s->next_id++;
        q->t_submit = Clock::now();
        s->pending.push_back(q);
        s->reqs[q->id] = q;
        s->st.submitted++;
        *ticket = q->id;
        lk.unlock();
        s->cv_sched.notify_one();
        return TRID_OK;
    }
    if (req->kind == TRID_REQ_OPERATOR) {
        auto it = s->ops.find(req->op);
        if (it == s->ops.end()) {
            delete q;
            return TRID_EINVAL;
        }
        q->sizes = it->second.sizes;
        q->unknowns = it->second.unknowns;
    } else {
        if (!req->a || !req->b || !req->c || !req->offsets || req->num_systems < 1) {
            delete q;
            return TRID_EINVAL;
        }
        for (int64_t i = 0; i < req->num_systems; ++i) {
            const int64_t n = req->offsets[i + 1] - req->offsets[i];
            if (n < 2 || n > trid::kRaggedMaxSize) {
                delete q;
                return n < 2 ? TRID_EINVAL : TRID_ENOTSUP;
            }
            q->sizes.push_back(n);
            q->unknowns += n;
        }
        // offsets are relative to the request's arrays
        if (req->offsets[0] != 0) {
            q->r.a += req->offsets[0];
            q->r.b += req->offsets[0];
            q->r.c += req->offsets[0];
            q->r.d += req->offsets[0];
            q->r.x += req->offsets[0];
        }
    }
    // a request that could never be admitted is rejected now
    const BatchLayout L(q->unknowns, (int64_t)q->sizes.size());
    const size_t cap = (size_t)s->o.hbm_budget_bytes > s->ops_bytes
                           ? (size_t)s->o.hbm_budget_bytes - s->ops_bytes
                           : 0;
    if (q->unknowns > s->o.max_batch_unknowns ||
        (int64_t)q->sizes.size() > s->o.max_batch_systems || L.bytes > cap) {
        delete q;
        return TRID_ENOMEM;
    }
    if (q->r.op >= 0) s->ops[q->r.op].refs++;
    q->id = s->next_id++;
    q->t_submit = Clock::now();
    s->pending.push_back(q);
    s->reqs[q->id] = q;
    s->st.submitted++;
    *ticket = q->id;
    lk.unlock();
    s->cv_sched.notify_one();
    return TRID_OK;
}

int trid_server_wait(trid_server s, uint64_t ticket, int64_t timeout_us, trid_request_result* res)
{
    if (!s) return TRID_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return TRID_EINVAL;
    Req* q = it->second;
    auto final_ = [q] { return q->state >= kDone; };
    // several threads may wait on one ticket: the last one out frees it
    q->waiters++;
    bool done = true;
    if (timeout_us < 0) s->cv_done.wait(lk, final_);
    else done = s->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), final_);
    q->waiters--;
    if (!done) return TRID_EAGAIN;
    if (res) {
        res->status = q->status;
        res->state = q->state;
        res->latency_us = us_between(q->t_submit, q->t_done);
        res->queue_us = q->state == kCancelled ? res->latency_us
                                               : us_between(q->t_submit, q->t_launch);
        res->batch = q->batch;
        res->batch_systems = q->batch_systems;
    }
    if (!q->retired) {
        q->retired = true;
        s->reqs.erase(ticket);
    }
    if (q->waiters == 0) delete q;
    return TRID_OK;
}

int trid_server_cancel(trid_server s, uint64_t ticket)
{
    if (!s) return TRID_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end() || it->second->state != kQueued) return TRID_EINVAL;
    Req* q = it->second;
    s->pending.erase(std::find(s->pending.begin(), s->pending.end(), q));
    s->finish(q, kCancelled, TRID_OK);
    s->cv_done.notify_all();
    return TRID_OK;
}

int trid_server_get_stats(trid_server s, trid_server_stats* out)
{
    if (!s || !out) return TRID_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->queued = (int64_t)s->pending.size();
    out->device_bytes_in_use = (int64_t)(s->in_use + s->ops_bytes);
    return TRID_OK;
}

}  // extern "C"
