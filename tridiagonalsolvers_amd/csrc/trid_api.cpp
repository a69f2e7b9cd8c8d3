// trid_api.cpp -- the C ABI of libtrid.so (declared in include/trid.h).
//
// Validation, planning (which kernel family and launch shape serves a
// request), the stream-ordered device API and the host-pointer drop-ins that
// replace the reference wrappers (cr_app.cu:4, pcr_app.cu:5, crpcr_app.cu:5,
// crNBC.cu:123,344, crNBC32banks.cu:123, pcrGauss_app.cu:164).
//
// Every entry point returns a trid_status and records it as the calling
// thread's last status; nothing here exits the process or prints.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>

#include "trid.h"
#include "trid_launch.hpp"
#include "trid_plan.hpp"

struct trid_context {
    int device;
    void* ws;         // workspace: two-pass pcrGauss, separator form, interleaved, pivoting
    size_t ws_bytes;
    bool fixed;       // ws is the library's fixed scratch array: never grown or freed
};

namespace {

thread_local int g_last_status = TRID_OK;

// Grow the handle's workspace to at least `bytes` (device memory; freeing
// the old buffer waits for work that may still use it).
int ensure_ws(trid_context* h, size_t bytes)
{
    if (!h) return TRID_EINVAL;
    if (h->ws_bytes >= bytes) return TRID_OK;
    if (h->fixed) return TRID_ENOMEM;  // the no-allocation entry point never allocates
    if (h->ws) {
        (void)hipFree(h->ws);
        h->ws = nullptr;
        h->ws_bytes = 0;
    }
    hipError_t e = hipMalloc(&h->ws, bytes);
    if (e != hipSuccess) {
        h->ws = nullptr;
        return e == hipErrorOutOfMemory ? TRID_ENOMEM : TRID_EHIP;
    }
    h->ws_bytes = bytes;
    return TRID_OK;
}

int set_status(int s)
{
    g_last_status = s;
    return s;
}

int from_hip(hipError_t e)
{
    if (e == hipSuccess) return TRID_OK;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return TRID_ENOMEM;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return TRID_ENODEV;
    return TRID_EHIP;
}

// Only gfx950 code objects are built; check the current device once.
std::mutex g_dev_mu;
int g_dev_checked[64];  // 0 unknown, 1 ok, 2 not gfx950

int check_device()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return TRID_ENODEV;
    if (dev < 0 || dev >= 64) return TRID_ENODEV;
    std::lock_guard<std::mutex> lk(g_dev_mu);
    if (g_dev_checked[dev] == 0) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return TRID_ENODEV;
        g_dev_checked[dev] = (std::strncmp(prop.gcnArchName, "gfx950", 6) == 0) ? 1 : 2;
    }
    return g_dev_checked[dev] == 1 ? TRID_OK : TRID_ENODEV;
}

bool aligned16(const void* p)
{
    return (reinterpret_cast<uintptr_t>(p) & 15u) == 0;
}

// Device workspace the plan's kernels need (0 for the register, LDS and
// in-LDS pcrGauss forms).
template <typename T>
size_t plan_ws_bytes(const trid::Plan& p, int method, int64_t n, int64_t N)
{
    if (p.family == trid::Family::PcrThomas && p.two_pass)
        return (size_t)trid::pcrg_ws_elems(n * N) * sizeof(T);
    if (p.family == trid::Family::Part)
        return trid::part_ws_bytes<T>(method == TRID_CR ? TRID_CR : TRID_CRPCR, n, N);
    return 0;
}

template <typename T>
int solve_impl(trid_handle h, int method, const T* a, const T* b, const T* c, const T* d, T* x,
               int64_t n, int64_t N, const trid_opts* opts, hipStream_t stream)
{
    trid_opts o{};
    if (opts) o = *opts;
    if (!h) return TRID_EINVAL;
    if (!a || !b || !c || !d || !x) return TRID_EINVAL;
    if (n < 2 || N < 0) return TRID_EINVAL;
    if (o.flags & ~TRID_FLAG_TWO_PASS) return TRID_EINVAL;
    if (N == 0) return TRID_OK;

    trid::Plan p;
    int st = trid::make_plan(method, n, N, (int)sizeof(T), o, &p);
    if (st != TRID_OK) return st;
    st = check_device();
    if (st != TRID_OK) return st;

    switch (p.family) {
        case trid::Family::Reg: {
            trid::RegArgs<T> g{};
            g.a = a;
            g.b = b;
            g.c = c;
            g.d = d;
            g.x = x;
            g.n = n;
            g.N = N;
            g.M = p.M;
            g.L = p.L;
            g.rest = p.rest;
            g.io = p.io;
            g.vec = (p.dense && aligned16(a) && aligned16(b) && aligned16(c) && aligned16(d) &&
                     aligned16(x))
                        ? 1
                        : 0;
            g.stream = stream;
            return from_hip(g.L <= 64 ? trid::launch_reg_crpcr<T>(g) : trid::launch_reg_mw<T>(g));
        }
        case trid::Family::Lds: {
            trid::LdsArgs<T> g{};
            g.a = a;
            g.b = b;
            g.c = c;
            g.d = d;
            g.x = x;
            g.n = n;
            g.N = N;
            g.n_pad = (int)p.n_pad;
            g.rest = p.rest;
            g.pad = p.pad;
            g.stream = stream;
            return from_hip(trid::launch_lds_crpcr<T>(g));
        }
        case trid::Family::PcrThomas: {
            trid::PcrThomasArgs<T> g{};
            g.a = a;
            g.b = b;
            g.c = c;
            g.d = d;
            g.x = x;
            g.n = n;
            g.N = N;
            g.n_pad = (int)p.n_pad;
            g.G = p.gaussian;
            g.pad = p.pad;
            g.two_pass = p.two_pass;
            if (p.two_pass) {
                st = ensure_ws(h, plan_ws_bytes<T>(p, method, n, N));
                if (st != TRID_OK) return st;
                g.ws = static_cast<T*>(h->ws);
            }
            g.stream = stream;
            return from_hip(trid::launch_pcr_thomas<T>(g));
        }
        case trid::Family::Part: {
            const int m = method == TRID_CR ? TRID_CR : TRID_CRPCR;
            st = ensure_ws(h, plan_ws_bytes<T>(p, method, n, N));
            if (st != TRID_OK) return st;
            trid::PartArgs<T> g{};
            g.a = a;
            g.b = b;
            g.c = c;
            g.d = d;
            g.x = x;
            g.n = n;
            g.N = N;
            g.method = m;
            g.rest = p.rest;
            g.ws = h->ws;
            g.stream = stream;
            return from_hip(trid::launch_part<T>(g));
        }
        default:
            return TRID_ENOTSUP;
    }
}

// Strided batch: system s at s * bs (bs >= n): the register kernels, and the
// separator form for systems longer than they hold.
template <typename T>
int strided_impl(trid_context* h, int method, const T* a, const T* b, const T* c, const T* d,
                 T* x, int64_t n, int64_t N, int64_t bs, const trid_opts* opts,
                 hipStream_t stream)
{
    trid_opts o{};
    if (opts) o = *opts;
    if (!h || !a || !b || !c || !d || !x) return TRID_EINVAL;
    if (n < 2 || N < 0 || bs < n) return TRID_EINVAL;
    if (o.flags != 0) return TRID_EINVAL;
    if (method != TRID_CR && method != TRID_PCR && method != TRID_CRPCR && method != TRID_AUTO)
        return TRID_ENOTSUP;
    if (N == 0) return TRID_OK;
    if (N > 1 && bs > (INT64_MAX - n) / (N - 1)) return TRID_EINVAL;
    if (bs == n) return solve_impl<T>(h, method, a, b, c, d, x, n, N, &o, stream);
    trid::Plan p;
    int st = trid::make_plan(method, n, N, (int)sizeof(T), o, &p);
    if (st != TRID_OK) return st;
    if (p.family != trid::Family::Reg && p.family != trid::Family::Part) return TRID_ENOTSUP;
    st = check_device();
    if (st != TRID_OK) return st;
    if (p.family == trid::Family::Part) {
        const int m = method == TRID_CR ? TRID_CR : TRID_CRPCR;
        st = ensure_ws(h, trid::part_ws_bytes<T>(m, n, N));
        if (st != TRID_OK) return st;
        trid::PartArgs<T> g{};
        g.a = a;
        g.b = b;
        g.c = c;
        g.d = d;
        g.x = x;
        g.n = n;
        g.N = N;
        g.bs = bs;
        g.method = m;
        g.rest = p.rest;
        g.ws = h->ws;
        g.stream = stream;
        return from_hip(trid::launch_part<T>(g));
    }
    trid::RegArgs<T> g{};
    g.a = a;
    g.b = b;
    g.c = c;
    g.d = d;
    g.x = x;
    g.n = n;
    g.N = N;
    g.M = p.M;
    g.L = p.L;
    g.rest = p.rest;
    g.bs = bs;
    // 16-byte vectors need every system to start 16-byte aligned (the
    // wave-staged form then reads each system's 16-byte pieces in place)
    g.vec = (p.dense && (bs * (int64_t)sizeof(T)) % 16 == 0 && aligned16(a) && aligned16(b) &&
             aligned16(c) && aligned16(d) && aligned16(x))
                ? 1
                : 0;
    g.io = p.io;
    g.stream = stream;
    return from_hip(g.L <= 64 ? trid::launch_reg_crpcr<T>(g) : trid::launch_reg_mw<T>(g));
}

// ---- solve without device allocations -------------------------------------
// The workspace of trid_solve_noalloc: one fixed array per device, loaded
// with the library's code object and held until the process ends. In host
// code `scratch` names the runtime's host-side stand-in, not device memory;
// kernels get the device address that hipGetSymbolAddress reports for the
// current device (looked up on the first call on each device, kept here).
__device__ __align__(256) char scratch[TRID_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];

int scratch_ptr(void** out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return TRID_ENODEV;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return TRID_EHIP;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return TRID_OK;
}

// trid_solve_* with the workspace taken from `scratch`: the size check comes
// first (a request that needs more than the array fails with TRID_ENOMEM
// before anything touches the device), then the same kernels as the handle
// path run with the array as their workspace.
template <typename T>
int noalloc_impl(int method, const T* a, const T* b, const T* c, const T* d, T* x, int64_t n,
                 int64_t N, const trid_opts* opts, hipStream_t stream)
{
    trid_opts o{};
    if (opts) o = *opts;
    if (!a || !b || !c || !d || !x) return TRID_EINVAL;
    if (n < 2 || N < 0) return TRID_EINVAL;
    if (o.flags & ~TRID_FLAG_TWO_PASS) return TRID_EINVAL;
    if (N == 0) return TRID_OK;
    trid::Plan p;
    int st = trid::make_plan(method, n, N, (int)sizeof(T), o, &p);
    if (st != TRID_OK) return st;
    if (plan_ws_bytes<T>(p, method, n, N) > TRID_NOALLOC_SCRATCH_BYTES) return TRID_ENOMEM;
    st = check_device();
    if (st != TRID_OK) return st;
    trid_context ctx{};
    st = scratch_ptr(&ctx.ws);
    if (st != TRID_OK) return st;
    (void)hipGetDevice(&ctx.device);
    ctx.ws_bytes = TRID_NOALLOC_SCRATCH_BYTES;
    ctx.fixed = true;
    return solve_impl<T>(&ctx, method, a, b, c, d, x, n, N, &o, stream);
}

// Host-pointer drop-in: copy in, one solve on the null stream, copy out.
template <typename T>
int host_solve(int method, const T* a, const T* b, const T* c, const T* d, T* x, int n, int N,
               int rest, int gaussian)
{
    if (!a || !b || !c || !d || !x) return TRID_EINVAL;
    if (n < 2 || N < 0) return TRID_EINVAL;
    if (N == 0) return TRID_OK;
    trid_opts o{};
    o.rest = rest;
    o.gaussian = gaussian;
    trid::Plan p;
    int st = trid::make_plan(method, n, N, (int)sizeof(T), o, &p);
    if (st != TRID_OK) return st;
    st = check_device();
    if (st != TRID_OK) return st;

    const size_t elems = (size_t)n * (size_t)N;
    const size_t bytes = elems * sizeof(T);
    // one allocation, five 256-byte aligned arrays
    const size_t slot = (bytes + 255) & ~(size_t)255;
    char* base = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&base), slot * 5);
    if (e != hipSuccess) return from_hip(e);
    T* da = reinterpret_cast<T*>(base);
    T* db = reinterpret_cast<T*>(base + slot);
    T* dc = reinterpret_cast<T*>(base + 2 * slot);
    T* dd = reinterpret_cast<T*>(base + 3 * slot);
    T* dx = reinterpret_cast<T*>(base + 4 * slot);
    int rc = TRID_OK;
    if ((e = hipMemcpy(da, a, bytes, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(db, b, bytes, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(dc, c, bytes, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(dd, d, bytes, hipMemcpyHostToDevice)) != hipSuccess) {
        rc = from_hip(e);
    }
    trid_context ctx{};
    if (rc == TRID_OK) rc = solve_impl<T>(&ctx, method, da, db, dc, dd, dx, n, N, &o, nullptr);
    if (rc == TRID_OK) {
        e = hipStreamSynchronize(nullptr);
        if (e == hipSuccess) e = hipMemcpy(x, dx, bytes, hipMemcpyDeviceToHost);
        rc = from_hip(e);
    }
    if (ctx.ws) (void)hipFree(ctx.ws);
    (void)hipFree(base);
    return rc;
}

template <typename T>
int interleaved_impl(trid_context* h, const T* a, const T* b, const T* c, const T* d, T* x,
                     int64_t n, int64_t N, int64_t stride, hipStream_t stream)
{
    if (!h || !a || !b || !c || !d || !x) return TRID_EINVAL;
    if (n < 2 || N < 0) return TRID_EINVAL;
    if (x == a || x == b || x == c) return TRID_EINVAL;
    if (N == 0) return TRID_OK;
    if (stride < N) return TRID_EINVAL;
    if (n > INT64_MAX / stride / (int64_t)sizeof(T)) return TRID_EINVAL;
    int st = check_device();
    if (st != TRID_OK) return st;
    // the register form (256 < n <= 1024 in FP32, 128 < n <= 1024 in FP64)
    // needs no workspace; the two-ended elimination keeps its c' factors in
    // n * stride elements
    if (!trid::interleaved_reg_takes<T>(n, N)) {
        st = ensure_ws(h, (size_t)n * (size_t)stride * sizeof(T));
        if (st != TRID_OK) return st;
    }
    trid::InterleavedArgs<T> g{};
    g.a = a;
    g.b = b;
    g.c = c;
    g.d = d;
    g.x = x;
    g.n = n;
    g.N = N;
    g.stride = stride;
    g.ws = static_cast<T*>(h->ws);
    g.stream = stream;
    return from_hip(trid::launch_interleaved<T>(g));
}

// Partial-pivoting solve (LAPACK ?gtsv semantics). stride 0 = system-major.
template <typename T>
int pivot_impl(trid_context* h, const T* a, const T* b, const T* c, const T* d, T* x,
               int64_t n, int64_t N, int64_t stride, int32_t* info, hipStream_t stream)
{
    if (!h || !a || !b || !c || !d || !x) return TRID_EINVAL;
    if (n < 1 || N < 0 || stride < 0) return TRID_EINVAL;
    if (x == a || x == b || x == c) return TRID_EINVAL;
    if (N == 0) return TRID_OK;
    if (stride > 0) {
        if (stride < N) return TRID_EINVAL;
        if (n > INT64_MAX / stride / (int64_t)sizeof(T)) return TRID_EINVAL;
    } else if (n > INT64_MAX / N / (int64_t)sizeof(T)) {
        return TRID_EINVAL;
    }
    int st = check_device();
    if (st != TRID_OK) return st;
    st = ensure_ws(h, trid::pivot_ws_bytes<T>(n, N, stride > 0));
    if (st != TRID_OK) return st;
    trid::PivotArgs<T> g{};
    g.a = a;
    g.b = b;
    g.c = c;
    g.d = d;
    g.x = x;
    g.n = n;
    g.N = N;
    g.stride = stride;
    g.ws = static_cast<T*>(h->ws);
    g.info = info;
    g.stream = stream;
    return from_hip(trid::launch_pivot<T>(g));
}

// Host-pointer pivoting drop-in: copy in, solve on the null stream, copy out.
template <typename T>
int host_gtsv(const T* a, const T* b, const T* c, const T* d, T* x, int n, int N)
{
    if (!a || !b || !c || !d || !x) return TRID_EINVAL;
    if (n < 1 || N < 0) return TRID_EINVAL;
    if (N == 0) return TRID_OK;
    int st = check_device();
    if (st != TRID_OK) return st;
    const size_t bytes = (size_t)n * (size_t)N * sizeof(T);
    const size_t slot = (bytes + 255) & ~(size_t)255;
    char* base = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&base), slot * 5);
    if (e != hipSuccess) return from_hip(e);
    T* dv[5];
    for (int k = 0; k < 5; ++k) dv[k] = reinterpret_cast<T*>(base + k * slot);
    const T* hv[4] = {a, b, c, d};
    int rc = TRID_OK;
    for (int k = 0; k < 4 && rc == TRID_OK; ++k)
        rc = from_hip(hipMemcpy(dv[k], hv[k], bytes, hipMemcpyHostToDevice));
    trid_context ctx{};
    if (rc == TRID_OK)
        rc = pivot_impl<T>(&ctx, dv[0], dv[1], dv[2], dv[3], dv[4], n, N, 0, nullptr, nullptr);
    if (rc == TRID_OK) {
        e = hipStreamSynchronize(nullptr);
        if (e == hipSuccess) e = hipMemcpy(x, dv[4], bytes, hipMemcpyDeviceToHost);
        rc = from_hip(e);
    }
    if (ctx.ws) (void)hipFree(ctx.ws);
    (void)hipFree(base);
    return rc;
}

}  // namespace

extern "C" {

int trid_solve_interleaved_f32(trid_handle handle, const float* a, const float* b,
                               const float* c, const float* d, float* x, int64_t systemSize,
                               int64_t numSystems, int64_t batchStride, void* stream)
{
    return set_status(interleaved_impl<float>(handle, a, b, c, d, x, systemSize, numSystems,
                                              batchStride, static_cast<hipStream_t>(stream)));
}

int trid_solve_interleaved_f64(trid_handle handle, const double* a, const double* b,
                               const double* c, const double* d, double* x, int64_t systemSize,
                               int64_t numSystems, int64_t batchStride, void* stream)
{
    return set_status(interleaved_impl<double>(handle, a, b, c, d, x, systemSize, numSystems,
                                               batchStride, static_cast<hipStream_t>(stream)));
}

int trid_solve_strided_f32(trid_handle handle, int method, const float* a, const float* b,
                           const float* c, const float* d, float* x, int64_t systemSize,
                           int64_t numSystems, int64_t batchStride, const trid_opts* opts,
                           void* stream)
{
    return set_status(strided_impl<float>(handle, method, a, b, c, d, x, systemSize, numSystems,
                                          batchStride, opts, static_cast<hipStream_t>(stream)));
}

int trid_solve_strided_f64(trid_handle handle, int method, const double* a, const double* b,
                           const double* c, const double* d, double* x, int64_t systemSize,
                           int64_t numSystems, int64_t batchStride, const trid_opts* opts,
                           void* stream)
{
    return set_status(strided_impl<double>(handle, method, a, b, c, d, x, systemSize,
                                           numSystems, batchStride, opts,
                                           static_cast<hipStream_t>(stream)));
}

int trid_solve_pivoting_f32(trid_handle handle, const float* a, const float* b, const float* c,
                            const float* d, float* x, int64_t systemSize, int64_t numSystems,
                            int32_t* info, void* stream)
{
    return set_status(pivot_impl<float>(handle, a, b, c, d, x, systemSize, numSystems, 0, info,
                                         static_cast<hipStream_t>(stream)));
}

int trid_solve_pivoting_f64(trid_handle handle, const double* a, const double* b,
                            const double* c, const double* d, double* x, int64_t systemSize,
                            int64_t numSystems, int32_t* info, void* stream)
{
    return set_status(pivot_impl<double>(handle, a, b, c, d, x, systemSize, numSystems, 0, info,
                                          static_cast<hipStream_t>(stream)));
}

int trid_solve_pivoting_interleaved_f32(trid_handle handle, const float* a, const float* b,
                                        const float* c, const float* d, float* x,
                                        int64_t systemSize, int64_t numSystems,
                                        int64_t batchStride, int32_t* info, void* stream)
{
    if (batchStride < 1) return set_status(TRID_EINVAL);
    return set_status(pivot_impl<float>(handle, a, b, c, d, x, systemSize, numSystems,
                                         batchStride, info, static_cast<hipStream_t>(stream)));
}

int trid_solve_pivoting_interleaved_f64(trid_handle handle, const double* a, const double* b,
                                        const double* c, const double* d, double* x,
                                        int64_t systemSize, int64_t numSystems,
                                        int64_t batchStride, int32_t* info, void* stream)
{
    if (batchStride < 1) return set_status(TRID_EINVAL);
    return set_status(pivot_impl<double>(handle, a, b, c, d, x, systemSize, numSystems,
                                          batchStride, info, static_cast<hipStream_t>(stream)));
}

int trid_gtsv_f32(const float* a, const float* b, const float* c, const float* d, float* x,
                  int systemSize, int numSystems)
{
    return set_status(host_gtsv<float>(a, b, c, d, x, systemSize, numSystems));
}

int trid_gtsv_f64(const double* a, const double* b, const double* c, const double* d,
                  double* x, int systemSize, int numSystems)
{
    return set_status(host_gtsv<double>(a, b, c, d, x, systemSize, numSystems));
}

const char* trid_status_string(int status)
{
    switch (status) {
        case TRID_OK: return "TRID_OK: success";
        case TRID_EINVAL: return "TRID_EINVAL: invalid argument";
        case TRID_ENOMEM: return "TRID_ENOMEM: out of memory";
        case TRID_EHIP: return "TRID_EHIP: HIP runtime error";
        case TRID_ENOTSUP: return "TRID_ENOTSUP: not supported by this library";
        case TRID_ENODEV: return "TRID_ENODEV: no gfx950 device";
        case TRID_EAGAIN: return "TRID_EAGAIN: not finished, try again";
        default: return "unknown trid status";
    }
}

int trid_last_status(void) { return g_last_status; }

int trid_device_count(void)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) return 0;
    int good = 0;
    for (int i = 0; i < count; ++i) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, i) == hipSuccess &&
            std::strncmp(prop.gcnArchName, "gfx950", 6) == 0)
            ++good;
    }
    return good;
}

int trid_create(trid_handle* handle)
{
    if (!handle) return set_status(TRID_EINVAL);
    trid_context* h = new (std::nothrow) trid_context();
    if (!h) return set_status(TRID_ENOMEM);
    h->device = 0;
    h->ws = nullptr;
    h->ws_bytes = 0;
    (void)hipGetDevice(&h->device);
    *handle = h;
    return set_status(TRID_OK);
}

int trid_destroy(trid_handle handle)
{
    if (!handle) return set_status(TRID_EINVAL);
    if (handle->ws) (void)hipFree(handle->ws);
    delete handle;
    return set_status(TRID_OK);
}

int trid_solve_f32(trid_handle handle, int method, const float* a, const float* b,
                   const float* c, const float* d, float* x, int64_t systemSize,
                   int64_t numSystems, const trid_opts* opts, void* stream)
{
    return set_status(solve_impl<float>(handle, method, a, b, c, d, x, systemSize, numSystems,
                                        opts, static_cast<hipStream_t>(stream)));
}

int trid_solve_f64(trid_handle handle, int method, const double* a, const double* b,
                   const double* c, const double* d, double* x, int64_t systemSize,
                   int64_t numSystems, const trid_opts* opts, void* stream)
{
    return set_status(solve_impl<double>(handle, method, a, b, c, d, x, systemSize, numSystems,
                                         opts, static_cast<hipStream_t>(stream)));
}

int trid_solve_noalloc(int method, int elemSize, const void* a, const void* b, const void* c,
                       const void* d, void* x, int64_t systemSize, int64_t numSystems,
                       const trid_opts* opts, void* stream)
{
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (elemSize == 4)
        return set_status(noalloc_impl<float>(
            method, static_cast<const float*>(a), static_cast<const float*>(b),
            static_cast<const float*>(c), static_cast<const float*>(d), static_cast<float*>(x),
            systemSize, numSystems, opts, s));
    if (elemSize == 8)
        return set_status(noalloc_impl<double>(
            method, static_cast<const double*>(a), static_cast<const double*>(b),
            static_cast<const double*>(c), static_cast<const double*>(d),
            static_cast<double*>(x), systemSize, numSystems, opts, s));
    return set_status(TRID_EINVAL);
}

int trid_solve_refined_f64(trid_handle handle, int method, const double* a, const double* b,
                           const double* c, const double* d, double* x, int64_t systemSize,
                           int64_t numSystems, int iters, const trid_opts* opts, void* stream)
{
    trid_opts o{};
    if (opts) o = *opts;
    if (!handle || !a || !b || !c || !d || !x) return set_status(TRID_EINVAL);
    if (systemSize < 2 || numSystems < 0 || iters < 0 || iters > 8) return set_status(TRID_EINVAL);
    if (o.flags != 0) return set_status(TRID_EINVAL);
    if (numSystems == 0) return set_status(TRID_OK);
    trid::Plan p;
    int st = trid::make_refine_plan(method, systemSize, numSystems, o, &p);
    if (st != TRID_OK) return set_status(st);
    st = check_device();
    if (st != TRID_OK) return set_status(st);
    trid::RefineArgs g{};
    g.a = a;
    g.b = b;
    g.c = c;
    g.d = d;
    g.x = x;
    g.n = systemSize;
    g.N = numSystems;
    g.M = p.M;
    g.L = p.L;
    g.rest = p.rest;
    g.iters = iters;
    g.vec = (p.dense && aligned16(a) && aligned16(b) && aligned16(c) && aligned16(d)) ? 1 : 0;
    g.stream = static_cast<hipStream_t>(stream);
    return set_status(from_hip(trid::launch_refine(g)));
}

int trid_residual_f64(const double* a, const double* b, const double* c, const double* d,
                      const double* x, const int64_t* offsets, int64_t numSystems,
                      double* relres, void* stream)
{
    if (!a || !b || !c || !d || !x || !offsets || !relres || numSystems < 0)
        return set_status(TRID_EINVAL);
    int st = check_device();
    if (st != TRID_OK) return set_status(st);
    return set_status(from_hip(trid::launch_residual(a, b, c, d, x, offsets, numSystems, relres,
                                                     static_cast<hipStream_t>(stream))));
}

int trid_generate_f32(int kind, uint64_t seed, float* a, float* b, float* c, float* d,
                      int64_t systemSize, int64_t numSystems, void* stream)
{
    if (!a || !b || !c || !d || systemSize < 1 || numSystems < 0 || kind < 0 || kind > 4)
        return set_status(TRID_EINVAL);
    int st = check_device();
    if (st != TRID_OK) return set_status(st);
    return set_status(from_hip(trid::launch_generate<float>(
        kind, seed, a, b, c, d, systemSize, numSystems, static_cast<hipStream_t>(stream))));
}

int trid_generate_f64(int kind, uint64_t seed, double* a, double* b, double* c, double* d,
                      int64_t systemSize, int64_t numSystems, void* stream)
{
    if (!a || !b || !c || !d || systemSize < 1 || numSystems < 0 || kind < 0 || kind > 4)
        return set_status(TRID_EINVAL);
    int st = check_device();
    if (st != TRID_OK) return set_status(st);
    return set_status(from_hip(trid::launch_generate<double>(
        kind, seed, a, b, c, d, systemSize, numSystems, static_cast<hipStream_t>(stream))));
}

#define TRID_DROPIN(NAME, METHOD)                                                              \
    int NAME##_f32(const float* a, const float* b, const float* c, const float* d, float* x,   \
                   int systemSize, int numSystems)                                             \
    {                                                                                          \
        return set_status(host_solve<float>(METHOD, a, b, c, d, x, systemSize, numSystems, 0,  \
                                            0));                                               \
    }                                                                                          \
    int NAME##_f64(const double* a, const double* b, const double* c, const double* d,         \
                   double* x, int systemSize, int numSystems)                                  \
    {                                                                                          \
        return set_status(host_solve<double>(METHOD, a, b, c, d, x, systemSize, numSystems, 0, \
                                             0));                                              \
    }

TRID_DROPIN(trid_cr, TRID_CR)
TRID_DROPIN(trid_pcr, TRID_PCR)
TRID_DROPIN(trid_crnbc, TRID_CRNBC)
TRID_DROPIN(trid_crnbc32, TRID_CRNBC32)
#undef TRID_DROPIN

int trid_crpcr_f32(const float* a, const float* b, const float* c, const float* d, float* x,
                   int systemSize, int numSystems, int restSystemSize)
{
    if (restSystemSize <= 0) return set_status(TRID_EINVAL);
    return set_status(host_solve<float>(TRID_CRPCR, a, b, c, d, x, systemSize, numSystems,
                                        restSystemSize, 0));
}
int trid_crpcr_f64(const double* a, const double* b, const double* c, const double* d,
                   double* x, int systemSize, int numSystems, int restSystemSize)
{
    if (restSystemSize <= 0) return set_status(TRID_EINVAL);
    return set_status(host_solve<double>(TRID_CRPCR, a, b, c, d, x, systemSize, numSystems,
                                         restSystemSize, 0));
}
int trid_crpcrnbc_f32(const float* a, const float* b, const float* c, const float* d,
                      float* x, int systemSize, int numSystems, int restSystemSize)
{
    if (restSystemSize <= 0) return set_status(TRID_EINVAL);
    return set_status(host_solve<float>(TRID_CRPCRNBC, a, b, c, d, x, systemSize, numSystems,
                                        restSystemSize, 0));
}
int trid_crpcrnbc_f64(const double* a, const double* b, const double* c, const double* d,
                      double* x, int systemSize, int numSystems, int restSystemSize)
{
    if (restSystemSize <= 0) return set_status(TRID_EINVAL);
    return set_status(host_solve<double>(TRID_CRPCRNBC, a, b, c, d, x, systemSize, numSystems,
                                         restSystemSize, 0));
}
int trid_pcrgauss_f32(const float* a, const float* b, const float* c, const float* d, float* x,
                      int systemSize, int numSystems, int gaussian)
{
    if (gaussian <= 0) return set_status(TRID_EINVAL);
    return set_status(host_solve<float>(TRID_PCRGAUSS, a, b, c, d, x, systemSize, numSystems,
                                        0, gaussian));
}
int trid_pcrgauss_f64(const double* a, const double* b, const double* c, const double* d,
                      double* x, int systemSize, int numSystems, int gaussian)
{
    if (gaussian <= 0) return set_status(TRID_EINVAL);
    return set_status(host_solve<double>(TRID_PCRGAUSS, a, b, c, d, x, systemSize, numSystems,
                                         0, gaussian));
}

}  // extern "C"
