// Ragged batches: many independent 1-D complex transforms of different
// lengths, packed back to back (job i at offset sum of the earlier lengths).
//
// Three ways to run the same batch:
//   afft_ragged_exec        one launch of the persistent work-queue kernel
//                           (plus the counter reset kernel)
//   afft_ragged_exec_graph  the same pair captured once in a hipGraph and
//                           replayed
//   afft_ragged_exec_per_size  every job through the plan of its own size,
//                           one plan execution per job: the launch-per-size
//                           form the persistent kernel is measured against
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>

#include "afft_internal.hpp"
#include "kragged.hpp"

using namespace afft;

struct afft_ragged_s {
    afft_context ctx = nullptr;
    int prec = 0;
    int flags = 0;
    std::vector<int64_t> sizes;
    std::vector<int64_t> offs;
    int64_t total = 0;
    Plan tables;                  // owns the device tables
    RJob* jobs_dev = nullptr;
    int njobs = 0;
    int* counter = nullptr;
    void* scratch = nullptr;
    int64_t scratch_pts = 0;
    int grid = 0;
    hipStream_t stream = nullptr;
    std::map<std::tuple<int, void*, const void*, void*>, hipGraphExec_t> graphs;
    std::map<int64_t, afft_plan> per_size;
    std::mutex mu;
    ~afft_ragged_s();
};

afft_ragged_s::~afft_ragged_s() {
    if (ctx) (void)hipSetDevice(ctx->device);
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    for (auto& kv : per_size) afft_plan_destroy(kv.second);
    if (jobs_dev) (void)hipFree(jobs_dev);
    if (counter) (void)hipFree(counter);
    if (scratch) (void)hipFree(scratch);
    if (stream) (void)hipStreamDestroy(stream);
}

static int schedule_into(Plan& tab, int prec, int64_t m, RJob& j, std::map<int64_t, const void*>& cache) {
    std::vector<int> rad = smooth_schedule(m);
    if (rad.empty() || (int)rad.size() > RAG_MAXPASS)
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "no radix schedule for " + std::to_string(m));
    j.npass = (int)rad.size();
    int64_t ns = 1;
    int off = 0;
    std::vector<double> tw;
    for (size_t i = 0; i < rad.size(); ++i) {
        j.rad[i] = rad[i];
        j.ns[i] = (int)ns;
        j.twoff[i] = off;
        if (i > 0) {
            for (int64_t k = 0; k < ns; ++k)
                for (int r = 1; r < rad[i]; ++r) {
                    double re, im;
                    root_of_unity(k * r, ns * rad[i], &re, &im);
                    tw.push_back(re);
                    tw.push_back(im);
                }
            off += (int)(ns * (rad[i] - 1));
        }
        ns *= rad[i];
    }
    auto it = cache.find(m);
    if (it != cache.end()) {
        j.tw = it->second;
        return AFFT_OK;
    }
    int rc;
    if (prec == 0) {
        std::vector<float> f(tw.begin(), tw.end());
        rc = upload_table(tab, f.data(), f.size() * 4, &j.tw);
    } else {
        rc = upload_table(tab, tw.data(), tw.size() * 8, &j.tw);
    }
    if (rc) return rc;
    cache[m] = j.tw;
    return AFFT_OK;
}

static int ragged_build(afft_ragged_s& r) {
    const int es = r.prec ? 16 : 8;
    std::vector<RJob> jobs(r.sizes.size());
    std::vector<double> cost(r.sizes.size());
    std::map<int64_t, const void*> twcache;
    std::map<int64_t, BlueTables> bcache;
    int64_t off = 0;
    r.offs.resize(r.sizes.size());
    for (size_t i = 0; i < r.sizes.size(); ++i) {
        const int64_t n = r.sizes[i];
        if (n < 1) return set_error(AFFT_ERR_INVALID_LENGTH, "job length must be >= 1");
        if (n > (1 << 20)) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "ragged jobs are limited to 2^20 points");
        RJob& j = jobs[i];
        std::memset(&j, 0, sizeof(j));
        j.in_off = j.out_off = off;
        r.offs[i] = off;
        off += n;
        j.n = (int)n;
        const bool smooth = !smooth_schedule(n).empty() || n == 1;
        if (n == 1) {
            // a 1-point transform: one radix-1-free schedule is the copy; use
            // m = 1 with a single trivial pass
            j.m = 1;
            j.npass = 1;
            j.rad[0] = 1;
            j.ns[0] = 1;
            j.tw = nullptr;
        } else if (smooth) {
            j.m = (int)n;
            int rc = schedule_into(r.tables, r.prec, n, j, twcache);
            if (rc) return rc;
        } else {
            j.blue = 1;
            j.m = (int)next_pow2(2 * n - 1);
            int rc = schedule_into(r.tables, r.prec, j.m, j, twcache);
            if (rc) return rc;
            auto it = bcache.find(n);
            if (it == bcache.end()) {
                BlueTables bt;
                r.tables.desc.precision = r.prec;
                r.tables.desc.flags = r.flags;
                if ((rc = blue_tables(r.tables, n, j.m, 1.0, bt))) return rc;
                it = bcache.emplace(n, bt).first;
            }
            j.chirp = it->second.chirp;
            j.filt[0] = it->second.filt[0];
            j.filt[1] = it->second.filt[1];
        }
        const double m = j.m;
        cost[i] = (j.blue ? 2.0 : 1.0) * m * std::log2(std::max(2.0, m)) * (2 * m > RAG_LDS ? 3.0 : 1.0);
        if (2 * j.m > RAG_LDS) r.scratch_pts = std::max<int64_t>(r.scratch_pts, j.m);
    }
    r.total = off;
    // longest-processing-time order
    std::vector<size_t> order(jobs.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    std::vector<RJob> sorted;
    sorted.reserve(jobs.size());
    for (size_t i : order) sorted.push_back(jobs[i]);
    r.njobs = (int)sorted.size();
    AFFT_HIP(hipMalloc(&r.jobs_dev, std::max<size_t>(1, sorted.size()) * sizeof(RJob)));
    if (!sorted.empty()) AFFT_HIP(hipMemcpy(r.jobs_dev, sorted.data(), sorted.size() * sizeof(RJob), hipMemcpyHostToDevice));
    AFFT_HIP(hipMalloc(&r.counter, 64));
    // two blocks per CU: each block holds 64 KiB of LDS
    r.grid = 2 * r.ctx->props.multiProcessorCount;
    if (r.grid < 1) r.grid = 1;
    const size_t sbytes = (size_t)r.grid * 2 * std::max<int64_t>(1, r.scratch_pts) * es;
    AFFT_HIP(hipMalloc(&r.scratch, sbytes));
    return AFFT_OK;
}

static int ragged_launch(afft_ragged_s& r, int dir, void* dst, const void* src, hipStream_t s) {
    RagArgs a{};
    a.in = src;
    a.out = dst;
    a.jobs = r.jobs_dev;
    a.njobs = r.njobs;
    a.dir = dir;
    a.counter = r.counter;
    a.scratch = r.scratch;
    a.scratch_pts = std::max<int64_t>(1, r.scratch_pts);
    void* rargs[] = {&r.counter};
    AFFT_HIP(hipLaunchKernel((const void*)&k_rag_reset, dim3(1), dim3(64), rargs, 0, s));
    void* args[] = {&a};
    const void* fn = r.prec ? (const void*)&k_ragged<double> : (const void*)&k_ragged<float>;
    AFFT_HIP(hipLaunchKernel(fn, dim3(r.grid), dim3(RAG_THREADS), args, 0, s));
    return AFFT_OK;
}

extern "C" {

int afft_ragged_create(afft_context ctx, int precision, const int64_t* sizes, int64_t count, int flags,
                       afft_ragged* out) {
    if (!ctx || !out) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL handle");
    *out = nullptr;
    if (!sizes) return set_error(AFFT_ERR_NIL_SLICE, "sizes is nil");
    if (count < 1) return set_error(AFFT_ERR_INVALID_LENGTH, "count must be >= 1");
    if (precision != AFFT_COMPLEX64 && precision != AFFT_COMPLEX128)
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "unknown precision");
    AFFT_HIP(hipSetDevice(ctx->device));
    auto* r = new afft_ragged_s();
    r->ctx = ctx;
    r->prec = precision;
    r->flags = flags;
    r->tables.ctx = ctx;
    r->tables.desc.precision = precision;
    r->sizes.assign(sizes, sizes + count);
    int rc = ragged_build(*r);
    if (rc == AFFT_OK) {
        hipError_t e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
        if (e != hipSuccess) rc = set_error(AFFT_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    if (rc != AFFT_OK) {
        std::string keep = afft_last_error();
        delete r;
        set_error(rc, keep);
        return rc;
    }
    *out = r;
    clear_error();
    return AFFT_OK;
}

int64_t afft_ragged_total_len(afft_ragged r) { return r ? r->total : 0; }

int afft_ragged_offsets(afft_ragged r, int64_t* offsets, int64_t cap) {
    if (!r || !offsets) return set_error(AFFT_ERR_NIL_SLICE, "NULL argument");
    if (cap < (int64_t)r->offs.size()) return set_error(AFFT_ERR_LENGTH_MISMATCH, "offsets array too short");
    std::copy(r->offs.begin(), r->offs.end(), offsets);
    clear_error();
    return AFFT_OK;
}

int afft_ragged_exec(afft_ragged r, int direction, void* dst, const void* src, void* stream) {
    if (!r) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL handle");
    if (!dst || !src) return set_error(AFFT_ERR_NIL_SLICE, "dst or src is nil");
    if (direction != AFFT_FORWARD && direction != AFFT_INVERSE)
        return set_error(AFFT_ERR_INVALID_ARGUMENT, "direction must be -1 or +1");
    if (dst == src) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "ragged batches run out of place");
    std::lock_guard<std::mutex> l(r->mu);
    AFFT_HIP(hipSetDevice(r->ctx->device));
    int rc = ragged_launch(*r, direction, dst, src, stream ? (hipStream_t)stream : r->stream);
    if (rc == AFFT_OK) clear_error();
    return rc;
}

int afft_ragged_exec_graph(afft_ragged r, int direction, void* dst, const void* src, void* stream) {
    if (!r) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL handle");
    if (!dst || !src) return set_error(AFFT_ERR_NIL_SLICE, "dst or src is nil");
    if (direction != AFFT_FORWARD && direction != AFFT_INVERSE)
        return set_error(AFFT_ERR_INVALID_ARGUMENT, "direction must be -1 or +1");
    if (dst == src) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "ragged batches run out of place");
    std::lock_guard<std::mutex> l(r->mu);
    AFFT_HIP(hipSetDevice(r->ctx->device));
    hipStream_t s = stream ? (hipStream_t)stream : r->stream;
    auto key = std::make_tuple(direction, dst, src, (void*)s);
    auto it = r->graphs.find(key);
    if (it == r->graphs.end()) {
        // capture on the plan's own stream, replay on the caller's
        hipGraph_t g;
        AFFT_HIP(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
        int rc = ragged_launch(*r, direction, dst, src, r->stream);
        hipError_t e = hipStreamEndCapture(r->stream, &g);
        if (rc != AFFT_OK) return rc;
        if (e != hipSuccess) return set_error(AFFT_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        hipGraphExec_t ex;
        e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return set_error(AFFT_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        it = r->graphs.emplace(key, ex).first;
    }
    AFFT_HIP(hipGraphLaunch(it->second, s));
    clear_error();
    return AFFT_OK;
}

int afft_ragged_exec_per_size(afft_ragged r, int direction, void* dst, const void* src, void* stream) {
    if (!r) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL handle");
    if (!dst || !src) return set_error(AFFT_ERR_NIL_SLICE, "dst or src is nil");
    std::lock_guard<std::mutex> l(r->mu);
    AFFT_HIP(hipSetDevice(r->ctx->device));
    const int es = r->prec ? 16 : 8;
    for (size_t i = 0; i < r->sizes.size(); ++i) {
        const int64_t n = r->sizes[i];
        auto it = r->per_size.find(n);
        if (it == r->per_size.end()) {
            afft_plan_desc d{};
            d.kind = AFFT_C2C;
            d.precision = r->prec;
            d.rank = 1;
            d.dims[0] = n;
            d.batch = 1;
            d.flags = r->flags;
            afft_plan p = nullptr;
            int rc = afft_plan_create_desc(r->ctx, &d, &p);
            if (rc) return rc;
            it = r->per_size.emplace(n, p).first;
        }
        const int64_t o = r->offs[i] * es;
        int rc = afft_plan_exec_device(it->second, direction, (char*)dst + o, (const char*)src + o,
                                       stream ? stream : (void*)r->stream);
        if (rc) return rc;
    }
    clear_error();
    return AFFT_OK;
}

int afft_ragged_destroy(afft_ragged r) {
    delete r;
    clear_error();
    return AFFT_OK;
}

}  // extern "C"
