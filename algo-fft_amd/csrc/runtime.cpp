// C ABI of algofft_hip: errors, kernel registry, Backend / Context / Buffer /
// Stream / PlanImpl entry points (the Go interfaces of gpu/backend.go:7-51).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <thread>
#include <vector>

#include "afft_internal.hpp"

namespace afft {

// ---- errors -------------------------------------------------------------------
static thread_local std::string g_last_error;

int set_error(int status, const std::string& msg) {
    g_last_error = msg;
    return status;
}
void clear_error() { g_last_error.clear(); }

// ---- registry ------------------------------------------------------------------
static std::vector<KernelInfo>& registry() {
    static std::vector<KernelInfo> r;
    return r;
}
void register_kernel(const KernelInfo& k) { registry().push_back(k); }

const KernelInfo* find_kernel(int prec, int n, int var, int dir) {
    for (const KernelInfo& k : registry())
        if (k.prec == prec && k.n == n && k.var == var && k.dir == dir) return &k;
    return nullptr;
}

uint64_t Context::features() const {
    // "gfx950" -> 0x950: the architecture stands in for the reference's CPU
    // feature bitmask in wisdom keys
    uint64_t f = 0;
    const char* a = props.gcnArchName;
    if (a[0] == 'g' && a[1] == 'f' && a[2] == 'x')
        for (const char* q = a + 3; *q && *q != ':'; ++q) {
            int v = (*q >= '0' && *q <= '9') ? *q - '0' : (*q >= 'a' && *q <= 'f') ? *q - 'a' + 10 : -1;
            if (v < 0) break;
            f = f * 16 + (uint64_t)v;
        }
    return f;
}

int max_single_pass(int prec) {
    int m = 1;
    for (const KernelInfo& k : registry())
        if (k.prec == prec && k.var == VAR_ROW && k.n > m && (k.n & (k.n - 1)) == 0) m = k.n;
    return m;
}

// ---- roots of unity ---------------------------------------------------------
// cos/sin of θ = 2π m/L with m reduced in integer arithmetic and θ folded
// into [0, π/4] before an extended-precision cosl/sinl.
static void cossin2pi(int64_t m, int64_t L, long double* c, long double* s) {
    m %= L;
    if (m < 0) m += L;
    const int64_t m8 = 8 * m;  // L < 2^59
    if (m8 <= L) {
        const long double two_pi = 6.283185307179586476925286766559005768L;
        long double a = two_pi * (long double)m / (long double)L;
        *c = cosl(a);
        *s = sinl(a);
    } else if (m8 <= 2 * L) {  // θ = π/2 - φ, φ = 2π (L-4m)/(4L)
        long double c2, s2;
        cossin2pi(L - 4 * m, 4 * L, &c2, &s2);
        *c = s2;
        *s = c2;
    } else if (m8 <= 4 * L) {  // θ = π - φ, φ = 2π (L-2m)/(2L)
        long double c2, s2;
        cossin2pi(L - 2 * m, 2 * L, &c2, &s2);
        *c = -c2;
        *s = s2;
    } else {                   // θ = 2π - φ, φ = 2π (L-m)/L
        long double c2, s2;
        cossin2pi(L - m, L, &c2, &s2);
        *c = c2;
        *s = -s2;
    }
}

void root_of_unity(int64_t m, int64_t L, double* re, double* im) {
    long double c, s;
    cossin2pi(m, L, &c, &s);
    *re = (double)c;
    *im = (double)-s;
}

void root_of_unity_ld(int64_t m, int64_t L, long double* re, long double* im) {
    long double c, s;
    cossin2pi(m, L, &c, &s);
    *re = c;
    *im = -s;
}

}  // namespace afft

using namespace afft;

// ---- C ABI ---------------------------------------------------------------------

static void copy_str(char* dst, size_t cap, const char* src) {
    if (!cap) return;
    std::strncpy(dst, src ? src : "", cap - 1);
    dst[cap - 1] = 0;
}

extern "C" {

const char* afft_last_error(void) { return g_last_error.c_str(); }

const char* afft_status_string(int status) {
    switch (status) {
    case AFFT_OK: return "ok";
    case AFFT_ERR_NO_BACKEND: return "algofft/gpu: no backend registered";
    case AFFT_ERR_BACKEND_UNAVAILABLE: return "algofft/gpu: backend unavailable";
    case AFFT_ERR_NOT_IMPLEMENTED: return "algofft/gpu: not implemented";
    case AFFT_ERR_INVALID_LENGTH: return "algofft/gpu: invalid length";
    case AFFT_ERR_NIL_SLICE: return "algofft/gpu: nil slice";
    case AFFT_ERR_LENGTH_MISMATCH: return "algofft/gpu: length mismatch";
    case AFFT_ERR_INVALID_STRIDE: return "algo-fft: invalid stride";
    case AFFT_ERR_INVALID_SPECTRUM: return "algo-fft: invalid spectrum";
    case AFFT_ERR_DEVICE: return "algofft/gpu: HIP runtime error";
    case AFFT_ERR_INVALID_ARGUMENT: return "algofft/gpu: invalid argument";
    }
    return "unknown status";
}

int afft_backend_get_info(afft_backend_info* out) {
    if (!out) return set_error(AFFT_ERR_NIL_SLICE, "info pointer is NULL");
    copy_str(out->name, sizeof(out->name), "hip");
    int rt = 0;
    (void)hipRuntimeGetVersion(&rt);
    char ver[64];
    std::snprintf(ver, sizeof(ver), "0.1 (hip runtime %d)", rt);
    copy_str(out->version, sizeof(out->version), ver);
    copy_str(out->description, sizeof(out->description),
             "AMD Instinct MI355X (gfx950) FFT backend: LDS Stockham, four-step, mixed radix, Bluestein");
    clear_error();
    return AFFT_OK;
}

int afft_backend_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n > 0 ? 1 : 0;
}

int afft_device_count(int* count) {
    if (!count) return set_error(AFFT_ERR_NIL_SLICE, "count pointer is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *count = 0;
        return set_error(AFFT_ERR_BACKEND_UNAVAILABLE, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
    }
    *count = n;
    clear_error();
    return AFFT_OK;
}

static int fill_device_info(int index, afft_device_info* out) {
    hipDeviceProp_t p;
    AFFT_HIP(hipGetDeviceProperties(&p, index));
    copy_str(out->name, sizeof(out->name), p.name);
    copy_str(out->vendor, sizeof(out->vendor), "AMD");
    int drv = 0;
    (void)hipDriverGetVersion(&drv);
    char d[64];
    std::snprintf(d, sizeof(d), "hip %d", drv);
    copy_str(out->driver, sizeof(out->driver), d);
    out->memory_mb = (int64_t)(p.totalGlobalMem >> 20);
    copy_str(out->compute_cap, sizeof(out->compute_cap), p.gcnArchName);
    return AFFT_OK;
}

int afft_device_get_info(int index, afft_device_info* out) {
    if (!out) return set_error(AFFT_ERR_NIL_SLICE, "info pointer is NULL");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return set_error(AFFT_ERR_BACKEND_UNAVAILABLE, "no HIP device");
    }
    if (index < 0 || index >= n) return set_error(AFFT_ERR_INVALID_ARGUMENT, "device index out of range");
    int rc = fill_device_info(index, out);
    if (rc == AFFT_OK) clear_error();
    return rc;
}

int afft_context_create(int device_index, afft_context* out) {
    if (!out) return set_error(AFFT_ERR_NIL_SLICE, "context out pointer is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return set_error(AFFT_ERR_BACKEND_UNAVAILABLE, "no HIP device");
    }
    if (device_index < 0 || device_index >= n)
        return set_error(AFFT_ERR_INVALID_ARGUMENT, "device index " + std::to_string(device_index) + " out of range");
    auto* c = new afft_context_s();
    c->device = device_index;
    hipError_t e = hipGetDeviceProperties(&c->props, device_index);
    if (e != hipSuccess) {
        delete c;
        return set_error(AFFT_ERR_DEVICE, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
    }
    // Set-up of the allocation-free entry point: the device address of its
    // static scratch on this device.  A failure here is retried, and
    // reported, by that entry point's first call.
    void* scr = nullptr;
    (void)noalloc_scratch(device_index, &scr);
    *out = c;
    clear_error();
    return AFFT_OK;
}

int afft_context_device(afft_context ctx, afft_device_info* out) {
    if (!ctx || !out) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL handle");
    int rc = fill_device_info(ctx->device, out);
    if (rc == AFFT_OK) clear_error();
    return rc;
}

int afft_context_destroy(afft_context ctx) {
    delete ctx;
    clear_error();
    return AFFT_OK;
}

int afft_buffer_create(afft_context ctx, int64_t elem_count, int precision, afft_buffer* out) {
    if (!ctx || !out) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL handle");
    *out = nullptr;
    if (elem_count < 0) return set_error(AFFT_ERR_INVALID_LENGTH, "negative element count");
    if (precision != AFFT_COMPLEX64 && precision != AFFT_COMPLEX128)
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "unknown precision");
    AFFT_HIP(hipSetDevice(ctx->device));
    auto* b = new afft_buffer_s();
    b->ctx = ctx;
    b->len = elem_count;
    b->precision = precision;
    size_t bytes = (size_t)elem_count * (precision == AFFT_COMPLEX64 ? 8 : 16);
    if (bytes) {
        hipError_t e = hipMalloc(&b->ptr, bytes);
        if (e != hipSuccess) {
            delete b;
            return set_error(AFFT_ERR_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
    }
    *out = b;
    clear_error();
    return AFFT_OK;
}

int64_t afft_buffer_len(afft_buffer b) { return b ? b->len : 0; }
int afft_buffer_precision(afft_buffer b) { return b ? b->precision : 0; }
void* afft_buffer_device_ptr(afft_buffer b) { return b ? b->ptr : nullptr; }

int afft_buffer_upload(afft_buffer b, const void* host, int64_t host_len) {
    if (!b) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL buffer");
    if (!host) return set_error(AFFT_ERR_NIL_SLICE, "host slice is nil");
    if (host_len < b->len) return set_error(AFFT_ERR_LENGTH_MISMATCH, "host slice shorter than buffer");
    AFFT_HIP(hipSetDevice(b->ctx->device));
    if (b->len) AFFT_HIP(hipMemcpy(b->ptr, host, (size_t)b->len * (b->precision ? 16 : 8), hipMemcpyHostToDevice));
    clear_error();
    return AFFT_OK;
}

int afft_buffer_download(afft_buffer b, void* host, int64_t host_len) {
    if (!b) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL buffer");
    if (!host) return set_error(AFFT_ERR_NIL_SLICE, "host slice is nil");
    if (host_len < b->len) return set_error(AFFT_ERR_LENGTH_MISMATCH, "host slice shorter than buffer");
    AFFT_HIP(hipSetDevice(b->ctx->device));
    if (b->len) AFFT_HIP(hipMemcpy(host, b->ptr, (size_t)b->len * (b->precision ? 16 : 8), hipMemcpyDeviceToHost));
    clear_error();
    return AFFT_OK;
}

int afft_buffer_destroy(afft_buffer b) {
    if (b) {
        if (b->ptr) {
            (void)hipSetDevice(b->ctx->device);
            (void)hipFree(b->ptr);
        }
        delete b;
    }
    clear_error();
    return AFFT_OK;
}

int afft_host_alloc(afft_context ctx, int64_t bytes, void** out) {
    if (!ctx || !out) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (bytes < 1) return set_error(AFFT_ERR_INVALID_LENGTH, "bytes must be >= 1");
    AFFT_HIP(hipSetDevice(ctx->device));
    AFFT_HIP(hipHostMalloc(out, (size_t)bytes, hipHostMallocDefault));
    clear_error();
    return AFFT_OK;
}

int afft_host_free(void* ptr) {
    if (!ptr) return set_error(AFFT_ERR_NIL_SLICE, "nil pointer");
    AFFT_HIP(hipHostFree(ptr));
    clear_error();
    return AFFT_OK;
}

int afft_stream_create(afft_context ctx, afft_stream* out) {
    if (!ctx || !out) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL handle");
    *out = nullptr;
    AFFT_HIP(hipSetDevice(ctx->device));
    auto* s = new afft_stream_s();
    s->ctx = ctx;
    hipError_t e = hipStreamCreateWithFlags(&s->s, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete s;
        return set_error(AFFT_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    *out = s;
    clear_error();
    return AFFT_OK;
}

int afft_stream_synchronize(afft_stream s) {
    if (!s) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL stream");
    AFFT_HIP(hipStreamSynchronize(s->s));
    clear_error();
    return AFFT_OK;
}

int afft_stream_destroy(afft_stream s) {
    if (s) {
        if (s->s) (void)hipStreamDestroy(s->s);
        delete s;
    }
    clear_error();
    return AFFT_OK;
}

void* afft_stream_native(afft_stream s) { return s ? (void*)s->s : nullptr; }

static int plan_common(afft_context ctx, const afft_plan_desc* d, afft_plan* out) {
    if (!ctx || !out) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL handle");
    *out = nullptr;
    if (!d) return set_error(AFFT_ERR_NIL_SLICE, "descriptor is NULL");
    AFFT_HIP(hipSetDevice(ctx->device));
    auto* p = new afft_plan_s();
    p->ctx = ctx;
    p->desc = *d;
    int rc = build_plan(*p);
    if (rc != AFFT_OK) {
        std::string keep = g_last_error;
        delete p;
        g_last_error = keep;
        return rc;
    }
    hipError_t e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete p;
        return set_error(AFFT_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    *out = p;
    clear_error();
    return AFFT_OK;
}

int afft_plan_create(afft_context ctx, int64_t n, int precision, const afft_plan_options* opts, afft_plan* out) {
    // gpu/plan.go:18-20: n < 1 is ErrInvalidLength
    if (n < 1) {
        if (out) *out = nullptr;
        return set_error(AFFT_ERR_INVALID_LENGTH, "n must be >= 1");
    }
    afft_plan_desc d{};
    d.kind = AFFT_C2C;
    d.precision = precision;
    d.rank = 1;
    d.dims[0] = n;
    d.batch = 1;
    if (opts && opts->in_place) d.flags |= AFFT_FLAG_INPLACE;
    return plan_common(ctx, &d, out);
}

int afft_plan_create_desc(afft_context ctx, const afft_plan_desc* desc, afft_plan* out) {
    return plan_common(ctx, desc, out);
}

int64_t afft_plan_len(afft_plan p) { return p ? p->n : 0; }
int afft_plan_precision(afft_plan p) { return p ? p->desc.precision : 0; }

int afft_plan_io_len(afft_plan p, int direction, int64_t* in_len, int64_t* out_len) {
    if (!p || !in_len || !out_len) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL argument");
    int di = direction == AFFT_INVERSE ? 1 : 0;
    *in_len = p->in_elems[di];
    *out_len = p->out_elems[di];
    clear_error();
    return AFFT_OK;
}

// ---- host-slice execution ------------------------------------------------------
// A batched 1-D plan whose host slices exceed kPipeMinBytes runs in chunks of
// whole transforms through kPipeSlots device slots on three streams: one
// queue of uploads, one of transforms, one of downloads, ordered by events.
// Chunk k's upload overlaps chunk k-1's transform and chunk k-2's download,
// so the host-to-device and device-to-host DMA engines both stay busy (the
// link is full duplex).  Pinned caller memory (afft_host_alloc or any
// hipHostMalloc block) is transferred directly; pageable memory goes through
// pinned staging slots, copied on the host while other slots' transfers run.
constexpr size_t kPipeMinBytes = size_t(64) << 20;
constexpr size_t kPipeChunkBytes = size_t(16) << 20;
constexpr int kPipeSlots = 3;

}  // extern "C"

namespace afft {
struct HostPipe {
    int dir_index = -1;
    int64_t chunk = 0;                 // transforms per full chunk
    afft_plan full = nullptr;          // plan of `chunk` transforms
    afft_plan tail = nullptr;
    int64_t tail_batch = 0;
    void* hin[kPipeSlots] = {};
    void* hout[kPipeSlots] = {};
    void* din[kPipeSlots] = {};
    void* dout[kPipeSlots] = {};
    hipStream_t up = nullptr, comp = nullptr, down = nullptr;
    hipEvent_t ev_up[kPipeSlots] = {}, ev_comp[kPipeSlots] = {}, ev_down[kPipeSlots] = {};
};

void destroy_host_pipe(HostPipe* h) {
    if (!h) return;
    for (hipStream_t st : {h->up, h->comp, h->down})
        if (st) (void)hipStreamSynchronize(st);
    for (int i = 0; i < kPipeSlots; ++i) {
        if (h->hin[i]) (void)hipHostFree(h->hin[i]);
        if (h->hout[i]) (void)hipHostFree(h->hout[i]);
        if (h->din[i]) (void)hipFree(h->din[i]);
        if (h->dout[i]) (void)hipFree(h->dout[i]);
        for (hipEvent_t ev : {h->ev_up[i], h->ev_comp[i], h->ev_down[i]})
            if (ev) (void)hipEventDestroy(ev);
    }
    for (hipStream_t st : {h->up, h->comp, h->down})
        if (st) (void)hipStreamDestroy(st);
    if (h->full) afft_plan_destroy(h->full);
    if (h->tail) afft_plan_destroy(h->tail);
    delete h;
}
}  // namespace afft

extern "C" {

// Host copy between a caller's pageable slice and a pinned staging slot.
// One core copies about 15-30 GB/s, below what the link moves each way, so
// copies of 8 MiB and more are split over four threads.
static void host_copy(void* dst, const void* src, size_t n) {
    constexpr size_t kSplitMin = size_t(8) << 20;
    constexpr int kThreads = 4;
    if (n < kSplitMin) {
        std::memcpy(dst, src, n);
        return;
    }
    const size_t part = (n / kThreads + 4095) & ~size_t(4095);
    std::vector<std::thread> ts;
    for (int t = 1; t < kThreads; ++t) {
        const size_t off = part * t;
        if (off >= n) break;
        const size_t len = std::min(part, n - off);
        ts.emplace_back([=] { std::memcpy((char*)dst + off, (const char*)src + off, len); });
    }
    std::memcpy(dst, src, std::min(part, n));
    for (auto& t : ts) t.join();
}

static bool host_pinned(const void* ptr) {
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, ptr) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return at.type == hipMemoryTypeHost;
}

// Child plan of the same description with `batch` transforms.
static int child_plan(afft_plan p, int64_t batch, afft_plan* out) {
    afft_plan_desc d = p->desc;
    d.batch = batch;
    return afft_plan_create_desc(static_cast<afft_context>(p->ctx), &d, out);
}

static bool pipe_eligible(afft_plan p, int di) {
    const afft_plan_desc& d = p->desc;
    if (d.rank != 1 || d.batch < 4) return false;
    if ((d.istride && d.istride != 1) || (d.ostride && d.ostride != 1)) return false;
    if (p->in_elems[di] % d.batch || p->out_elems[di] % d.batch) return false;
    const size_t ib = (size_t)p->in_elems[di] * p->in_esize[di];
    const size_t ob = (size_t)p->out_elems[di] * p->out_esize[di];
    return std::max(ib, ob) >= kPipeMinBytes;
}

static int host_exec_pipelined_run(afft_plan p, int dir, void* dst, const void* src);

// On an error part of the way through, the transfers already queued still
// read the caller's source and write its destination: wait for them before
// returning, so the caller's slices are not touched after the call.
static int host_exec_pipelined(afft_plan p, int dir, void* dst, const void* src) {
    const int rc = host_exec_pipelined_run(p, dir, dst, src);
    if (rc != AFFT_OK && p->pipe) {
        const std::string keep = afft_last_error();
        for (hipStream_t st : {p->pipe->up, p->pipe->comp, p->pipe->down})
            if (st) (void)hipStreamSynchronize(st);
        set_error(rc, keep);
    }
    return rc;
}

static int host_exec_pipelined_run(afft_plan p, int dir, void* dst, const void* src) {
    const int di = dir == AFFT_INVERSE ? 1 : 0;
    const int64_t B = p->desc.batch;
    const size_t ipt = (size_t)(p->in_elems[di] / B) * p->in_esize[di];    // bytes per transform
    const size_t opt = (size_t)(p->out_elems[di] / B) * p->out_esize[di];
    if (p->pipe && p->pipe->dir_index != di) {
        destroy_host_pipe(p->pipe);
        p->pipe = nullptr;
    }
    if (!p->pipe) {
        auto* h = new HostPipe();
        h->dir_index = di;
        h->chunk = std::max<int64_t>(1, (int64_t)(kPipeChunkBytes / std::max(ipt, opt)));
        h->chunk = std::min<int64_t>(h->chunk, (B + 3) / 4);
        int rc = child_plan(p, h->chunk, &h->full);
        hipError_t e = hipSuccess;
        for (int i = 0; i < kPipeSlots && !rc && e == hipSuccess; ++i) {
            e = hipHostMalloc(&h->hin[i], h->chunk * ipt, hipHostMallocDefault);
            if (e == hipSuccess) e = hipHostMalloc(&h->hout[i], h->chunk * opt, hipHostMallocDefault);
            if (e == hipSuccess) e = hipMalloc(&h->din[i], h->chunk * ipt);
            if (e == hipSuccess) e = hipMalloc(&h->dout[i], h->chunk * opt);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_up[i], hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_comp[i], hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_down[i], hipEventDisableTiming);
        }
        for (hipStream_t* st : {&h->up, &h->comp, &h->down})
            if (!rc && e == hipSuccess) e = hipStreamCreateWithFlags(st, hipStreamNonBlocking);
        if (!rc && e != hipSuccess) rc = set_error(AFFT_ERR_DEVICE, std::string("host pipeline: ") + hipGetErrorString(e));
        if (rc) {
            destroy_host_pipe(h);
            return rc;
        }
        p->pipe = h;
    }
    HostPipe& h = *p->pipe;
    const int64_t tail = B % h.chunk;
    if (tail && h.tail_batch != tail) {
        if (h.tail) afft_plan_destroy(h.tail);
        h.tail = nullptr;
        int rc = child_plan(p, tail, &h.tail);
        if (rc) return rc;
        h.tail_batch = tail;
    }
    const bool pin_in = host_pinned(src), pin_out = host_pinned(dst);
    const int64_t nch = (B + h.chunk - 1) / h.chunk;
    auto count_of = [&](int64_t k) { return std::min<int64_t>(h.chunk, B - k * h.chunk); };
    // chunk k's output leaves its pinned slot once its download has finished
    auto drain = [&](int64_t k) -> int {
        const int slot = (int)(k % kPipeSlots);
        AFFT_HIP(hipEventSynchronize(h.ev_down[slot]));
        if (!pin_out) host_copy((char*)dst + (size_t)(k * h.chunk) * opt, h.hout[slot], (size_t)count_of(k) * opt);
        return AFFT_OK;
    };
    for (int64_t k = 0; k < nch; ++k) {
        const int slot = (int)(k % kPipeSlots);
        const bool reuse = k >= kPipeSlots;   // the slot held chunk k - kPipeSlots
        if (reuse && !pin_out) {
            int rc = drain(k - kPipeSlots);
            if (rc) return rc;
        }
        const int64_t nb = count_of(k);
        afft_plan cp = nb == h.chunk ? h.full : h.tail;
        const char* sp = (const char*)src + (size_t)(k * h.chunk) * ipt;
        char* dp = (char*)dst + (size_t)(k * h.chunk) * opt;
        const void* upsrc = sp;
        if (!pin_in) {
            // the staging slot's previous upload must have finished reading it
            if (reuse) AFFT_HIP(hipEventSynchronize(h.ev_up[slot]));
            host_copy(h.hin[slot], sp, (size_t)nb * ipt);
            upsrc = h.hin[slot];
        }
        // upload into din[slot] once the slot's previous transform has read it
        if (reuse) AFFT_HIP(hipStreamWaitEvent(h.up, h.ev_comp[slot], 0));
        AFFT_HIP(hipMemcpyAsync(h.din[slot], upsrc, (size_t)nb * ipt, hipMemcpyHostToDevice, h.up));
        AFFT_HIP(hipEventRecord(h.ev_up[slot], h.up));
        // transform into dout[slot] once uploaded and once the slot's previous
        // download has read dout[slot]
        AFFT_HIP(hipStreamWaitEvent(h.comp, h.ev_up[slot], 0));
        if (reuse) AFFT_HIP(hipStreamWaitEvent(h.comp, h.ev_down[slot], 0));
        int rc = exec_plan(*cp, dir, h.dout[slot], h.din[slot], h.comp);
        if (rc) return rc;
        AFFT_HIP(hipEventRecord(h.ev_comp[slot], h.comp));
        AFFT_HIP(hipStreamWaitEvent(h.down, h.ev_comp[slot], 0));
        AFFT_HIP(hipMemcpyAsync(pin_out ? (void*)dp : h.hout[slot], h.dout[slot], (size_t)nb * opt,
                                hipMemcpyDeviceToHost, h.down));
        AFFT_HIP(hipEventRecord(h.ev_down[slot], h.down));
    }
    if (pin_out) {
        AFFT_HIP(hipStreamSynchronize(h.down));
    } else {
        for (int64_t k = std::max<int64_t>(0, nch - kPipeSlots); k < nch; ++k) {
            int rc = drain(k);
            if (rc) return rc;
        }
    }
    AFFT_HIP(hipStreamSynchronize(h.up));
    clear_error();
    return AFFT_OK;
}

static int host_exec(afft_plan p, int dir, void* dst, int64_t dst_len, const void* src, int64_t src_len) {
    if (!p) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "nil plan");
    if (!dst || !src) return set_error(AFFT_ERR_NIL_SLICE, "dst or src is nil");
    const int di = dir == AFFT_INVERSE ? 1 : 0;
    if (p->steps[di].empty()) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "direction not available for this plan kind");
    if (src_len < p->in_elems[di] || dst_len < p->out_elems[di])
        return set_error(AFFT_ERR_LENGTH_MISMATCH, "slice shorter than the plan requires");
    // PlanRealT.inverseSingle (plan_real_generic.go:298-316): DC and Nyquist
    // bins must be real to 1e-4 (complex64) / 1e-12 (complex128)
    if (di == 1 && p->desc.rank == 1 && (p->desc.kind == AFFT_R2C || p->desc.kind == AFFT_C2R)) {
        const int64_t h = p->desc.dims[0] / 2;
        const int64_t cd = p->desc.odist ? p->desc.odist : h + 1;
        const double eps = p->desc.precision ? 1e-12 : 1e-4;
        for (int64_t b = 0; b < p->desc.batch; ++b) {
            double i0, ih;
            if (p->desc.precision) {
                const double* x = (const double*)src + 2 * b * cd;
                i0 = x[1];
                ih = x[2 * h + 1];
            } else {
                const float* x = (const float*)src + 2 * b * cd;
                i0 = x[1];
                ih = x[2 * h + 1];
            }
            if (std::fabs(i0) > eps || std::fabs(ih) > eps)
                return set_error(AFFT_ERR_INVALID_SPECTRUM, "DC or Nyquist bin has a non-zero imaginary part");
        }
    }
    std::lock_guard<std::mutex> lock(p->mu);
    AFFT_HIP(hipSetDevice(p->ctx->device));
    if (pipe_eligible(p, di)) return host_exec_pipelined(p, dir, dst, src);
    size_t ib = (size_t)p->in_elems[di] * p->in_esize[di];
    size_t ob = (size_t)p->out_elems[di] * p->out_esize[di];
    if (p->stage_in_bytes < ib) {
        if (p->stage_in) (void)hipFree(p->stage_in);
        p->stage_in = nullptr;
        p->stage_in_bytes = 0;
        AFFT_HIP(hipMalloc(&p->stage_in, ib));
        p->stage_in_bytes = ib;
    }
    if (p->stage_out_bytes < ob) {
        if (p->stage_out) (void)hipFree(p->stage_out);
        p->stage_out = nullptr;
        p->stage_out_bytes = 0;
        AFFT_HIP(hipMalloc(&p->stage_out, ob));
        p->stage_out_bytes = ob;
    }
    AFFT_HIP(hipMemcpyAsync(p->stage_in, src, ib, hipMemcpyHostToDevice, p->stream));
    int rc = exec_plan(*p, dir, p->stage_out, p->stage_in, p->stream);
    if (rc != AFFT_OK) return rc;
    AFFT_HIP(hipMemcpyAsync(dst, p->stage_out, ob, hipMemcpyDeviceToHost, p->stream));
    AFFT_HIP(hipStreamSynchronize(p->stream));
    clear_error();
    return AFFT_OK;
}

int afft_plan_forward(afft_plan p, void* dst, int64_t dst_len, const void* src, int64_t src_len) {
    return host_exec(p, AFFT_FORWARD, dst, dst_len, src, src_len);
}

int afft_plan_inverse(afft_plan p, void* dst, int64_t dst_len, const void* src, int64_t src_len) {
    return host_exec(p, AFFT_INVERSE, dst, dst_len, src, src_len);
}

int afft_plan_exec_device(afft_plan p, int direction, void* dst, const void* src, void* stream) {
    if (!p) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "nil plan");
    if (!dst || !src) return set_error(AFFT_ERR_NIL_SLICE, "dst or src is nil");
    if (direction != AFFT_FORWARD && direction != AFFT_INVERSE)
        return set_error(AFFT_ERR_INVALID_ARGUMENT, "direction must be -1 or +1");
    if (p->steps[direction == AFFT_INVERSE ? 1 : 0].empty())
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "direction not available for this plan kind");
    std::lock_guard<std::mutex> lock(p->mu);
    AFFT_HIP(hipSetDevice(p->ctx->device));
    int rc = exec_plan(*p, direction, dst, src, stream ? (hipStream_t)stream : p->stream);
    if (rc == AFFT_OK) clear_error();
    return rc;
}

int afft_plan_exec_buffers(afft_plan p, int direction, afft_buffer dst, afft_buffer src, afft_stream s) {
    if (!p) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "nil plan");
    if (!dst || !src) return set_error(AFFT_ERR_NIL_SLICE, "dst or src buffer is nil");
    const int di = direction == AFFT_INVERSE ? 1 : 0;
    // Buffer lengths are in complex elements of the buffer's precision.
    const int ce = p->desc.precision ? 16 : 8;
    if (src->len * ce < p->in_elems[di] * p->in_esize[di] || dst->len * ce < p->out_elems[di] * p->out_esize[di])
        return set_error(AFFT_ERR_LENGTH_MISMATCH, "buffer shorter than the plan requires");
    if (src->precision != p->desc.precision || dst->precision != p->desc.precision)
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "buffer precision differs from plan precision");
    return afft_plan_exec_device(p, direction, dst->ptr, src->ptr, s ? (void*)s->s : nullptr);
}

int afft_plan_synchronize(afft_plan p) {
    if (!p) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL plan");
    AFFT_HIP(hipSetDevice(p->ctx->device));
    AFFT_HIP(hipStreamSynchronize(p->stream));
    clear_error();
    return AFFT_OK;
}

int afft_plan_describe(afft_plan p, char* buf, int cap) {
    if (!p || !buf || cap <= 0) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL argument");
    copy_str(buf, (size_t)cap, p->algo.c_str());
    clear_error();
    return AFFT_OK;
}

int64_t afft_plan_workspace_bytes(afft_plan p) {
    if (!p) return 0;
    int64_t t = 0;
    for (int i = 0; i < NTMP; ++i) t += (int64_t)p->tmp_bytes[i];
    return t;
}

int afft_plan_destroy(afft_plan p) {
    delete p;
    clear_error();
    return AFFT_OK;
}

}  // extern "C"
