// Ragged batches: many independent 1-D complex transforms of different
// lengths, packed back to back (job i at offset sum of the earlier lengths).
//
// Three ways to run the same batch:
//   afft_ragged_exec        the persistent work-queue kernels, one launch per
//                           task type (column tasks, row tasks, the
//                           Bluestein jobs' second column tasks, whole jobs;
//                           plain and Bluestein jobs apart), after one
//                           counter-reset kernel
//   afft_ragged_exec_graph  the same launches captured once in a hipGraph
//                           and replayed as one graph launch
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

constexpr int RAG_PHASES = RL_COUNT;

struct afft_ragged_s {
    afft_context ctx = nullptr;
    int prec = 0;
    int flags = 0;
    std::vector<int64_t> sizes;
    std::vector<int64_t> offs;
    int64_t total = 0;
    Plan tables;                  // owns the device tables
    RJob* jobs_dev = nullptr;
    RTask* tasks_dev = nullptr;
    int njobs = 0;
    int ntask[RAG_PHASES] = {};
    int task_off[RAG_PHASES] = {};
    int grid[2][RAG_PHASES] = {};       // [inverse][launch]
    int* counter = nullptr;
    void* scratch = nullptr;
    int64_t scratch_pts = 0;
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
    if (tasks_dev) (void)hipFree(tasks_dev);
    if (counter) (void)hipFree(counter);
    if (scratch) (void)hipFree(scratch);
    if (stream) (void)hipStreamDestroy(stream);
}

// Radix schedule and per-pass twiddles of length m (m <= RAG_WHOLE), shared
// between jobs through `cache`.
static int schedule_into(Plan& tab, int prec, int64_t m, RSched& j, std::map<int64_t, RSched>& cache) {
    auto it = cache.find(m);
    if (it != cache.end()) {
        j = it->second;
        return AFFT_OK;
    }
    std::memset(&j, 0, sizeof(j));
    if (m == 1) {
        j.npass = 1;
        j.rad[0] = 1;
        j.ns[0] = 1;
        j.nsm[0] = mod_magic(1);
        cache[m] = j;
        return AFFT_OK;
    }
    std::vector<int> rad = smooth_schedule(m);
    if (rad.empty() || (int)rad.size() > RAG_SPASS)
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "no radix schedule for " + std::to_string(m));
    j.npass = (int)rad.size();
    int64_t ns = 1;
    int off = 0;
    std::vector<double> tw;
    for (size_t i = 0; i < rad.size(); ++i) {
        j.rad[i] = rad[i];
        j.ns[i] = (int)ns;
        j.nsm[i] = mod_magic((unsigned)ns);
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
    int rc;
    if (tw.empty()) {
        j.tw = nullptr;
    } else if (prec == 0) {
        std::vector<float> f(tw.begin(), tw.end());
        if ((rc = upload_table(tab, f.data(), f.size() * 4, &j.tw))) return rc;
    } else {
        if ((rc = upload_table(tab, tw.data(), tw.size() * 8, &j.tw))) return rc;
    }
    cache[m] = j;
    return AFFT_OK;
}

// m = m1 * m2 with m1 <= m2 <= RAG_WHOLE and m1 as close to sqrt(m) as the
// divisors allow; false when m has no such split.
static bool split_len(int64_t m, int64_t* m1, int64_t* m2) {
    for (int64_t a = (int64_t)std::sqrt((double)m) + 1; a >= 2; --a) {
        if (a * a > m || m % a) continue;
        if (m / a > RAG_WHOLE) return false;
        *m1 = a;
        *m2 = m / a;
        return true;
    }
    return false;
}

static int pow2_floor(int64_t v) {
    int g = 1;
    while ((int64_t)g * 2 <= v) g *= 2;
    return g;
}

// The persistent kernel of launch p (RL_*) for direction dir.
template <typename T>
static const void* rag_kernel_t(int p, int dir) {
    switch (p) {
    case RL_COLS: return dir > 0 ? (const void*)&k_rag_cols<T, 1> : (const void*)&k_rag_cols<T, -1>;
    case RL_COLS_BLUE: return (const void*)&k_rag_cols<T, -1>;
    case RL_ROWS: return dir > 0 ? (const void*)&k_rag_rows<T, 1> : (const void*)&k_rag_rows<T, -1>;
    case RL_ROWS_BLUE: return (const void*)&k_rag_rows_blue<T>;
    case RL_COLS2: return (const void*)&k_rag_cols2<T>;
    case RL_WHOLE: return dir > 0 ? (const void*)&k_rag_whole<T, 1> : (const void*)&k_rag_whole<T, -1>;
    default: return (const void*)&k_rag_whole_blue<T>;
    }
}

static const void* rag_kernel(int prec, int p, int dir) {
    return prec ? rag_kernel_t<double>(p, dir) : rag_kernel_t<float>(p, dir);
}

static int ragged_build(afft_ragged_s& r) {
    const int es = r.prec ? 16 : 8;
    std::vector<RJob> jobs(r.sizes.size());
    std::map<int64_t, RSched> scache;
    std::map<std::pair<int64_t, int64_t>, BlueTables> bcache;
    struct TwSplit { int shift; const void* lo; const void* hi; };
    std::map<int64_t, TwSplit> tcache;
    struct Pending { RTask t; double cost; };
    std::vector<Pending> ph[RAG_PHASES];
    int64_t off = 0;
    r.offs.resize(r.sizes.size());
    r.scratch_pts = 0;
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
        int64_t m = n, m1 = 0, m2 = 0;
        bool blue = !(n == 1 || !smooth_schedule(n).empty());
        if (!blue && m > RAG_WHOLE && !split_len(m, &m1, &m2)) blue = true;   // no usable split
        if (blue) {
            m = next_pow2(2 * n - 1);
            if (m > RAG_WHOLE) split_len(m, &m1, &m2);
        }
        j.m = (int)m;
        j.blue = blue;
        const bool whole = m <= RAG_WHOLE;
        int rc;
        if (whole) {
            if ((rc = schedule_into(r.tables, r.prec, m, j.s[0], scache))) return rc;
        } else {
            j.m1 = (int)m1;
            j.m2 = (int)m2;
            if ((rc = schedule_into(r.tables, r.prec, m1, j.s[0], scache))) return rc;
            if ((rc = schedule_into(r.tables, r.prec, m2, j.s[1], scache))) return rc;
            auto tt = tcache.find(m);
            if (tt == tcache.end()) {
                TwSplit ts{};
                if ((rc = split_twiddles(r.tables, r.prec, m, &ts.shift, &ts.lo, &ts.hi))) return rc;
                tt = tcache.emplace(m, ts).first;
            }
            j.shift = tt->second.shift;
            j.twlo = tt->second.lo;
            j.twhi = tt->second.hi;
            // at most RAG_PTS points per task: every thread then holds at
            // most 16/R butterflies of a radix-R pass
            j.g_col = pow2_floor(std::min<int64_t>({RAG_THREADS, RAG_PTS / m1, m2}));
            j.g_row = pow2_floor(std::min<int64_t>({RAG_THREADS, RAG_PTS / m2, m1}));
            j.scr_off = r.scratch_pts;
            r.scratch_pts += m;
        }
        if (blue) {
            const auto key = std::make_pair(n, whole ? (int64_t)0 : m1);
            auto it = bcache.find(key);
            if (it == bcache.end()) {
                BlueTables bt;
                r.tables.desc.precision = r.prec;
                r.tables.desc.flags = r.flags;
                if ((rc = blue_tables(r.tables, n, m, 1.0, bt, key.second))) return rc;
                it = bcache.emplace(key, bt).first;
            }
            j.chirp = it->second.chirp;
            j.filt[0] = it->second.filt[0];
            j.filt[1] = it->second.filt[1];
        }
        const int job = (int)i;
        if (whole) {
            const double c = (blue ? 2.0 : 1.0) * m * std::log2(std::max<double>(2.0, (double)m));
            ph[blue ? RL_WHOLE_BLUE : RL_WHOLE].push_back({{job, RT_WHOLE, 0, 1}, c});
            continue;
        }
        // staging adds about one pass of LDS traffic to every task
        const double ccol = (double)m1 * (std::log2((double)m1) + 2.0);
        const double crow = (double)m2 * (std::log2((double)m2) + 2.0) * (blue ? 2.0 : 1.0);
        for (int64_t c0 = 0; c0 < m2; c0 += j.g_col) {
            const int cnt = (int)std::min<int64_t>(j.g_col, m2 - c0);
            ph[blue ? RL_COLS_BLUE : RL_COLS].push_back({{job, RT_COLS, (int)c0, cnt}, ccol * cnt});
            if (blue) ph[RL_COLS2].push_back({{job, RT_COLS2, (int)c0, cnt}, ccol * cnt});
        }
        for (int64_t r0 = 0; r0 < m1; r0 += j.g_row) {
            const int cnt = (int)std::min<int64_t>(j.g_row, m1 - r0);
            ph[blue ? RL_ROWS_BLUE : RL_ROWS].push_back({{job, RT_ROWS, (int)r0, cnt}, crow * cnt});
        }
    }
    r.total = off;
    // longest-processing-time order within each phase
    std::vector<RTask> tasks;
    for (int p = 0; p < RAG_PHASES; ++p) {
        std::stable_sort(ph[p].begin(), ph[p].end(), [](const Pending& a, const Pending& b) { return a.cost > b.cost; });
        r.task_off[p] = (int)tasks.size();
        r.ntask[p] = (int)ph[p].size();
        for (auto& q : ph[p]) tasks.push_back(q.t);
    }
    r.njobs = (int)jobs.size();
    AFFT_HIP(hipMalloc(&r.jobs_dev, std::max<size_t>(1, jobs.size()) * sizeof(RJob)));
    AFFT_HIP(hipMemcpy(r.jobs_dev, jobs.data(), jobs.size() * sizeof(RJob), hipMemcpyHostToDevice));
    AFFT_HIP(hipMalloc(&r.tasks_dev, std::max<size_t>(1, tasks.size()) * sizeof(RTask)));
    if (!tasks.empty())
        AFFT_HIP(hipMemcpy(r.tasks_dev, tasks.data(), tasks.size() * sizeof(RTask), hipMemcpyHostToDevice));
    AFFT_HIP(hipMalloc(&r.counter, 64));
    for (int inv = 0; inv < 2; ++inv)
        for (int p = 0; p < RAG_PHASES; ++p) {
            int per_cu = 0;
            AFFT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rag_kernel(r.prec, p, inv ? 1 : -1),
                                                                  RAG_THREADS, 0));
            r.grid[inv][p] = std::max(1, per_cu) * std::max(1, r.ctx->props.multiProcessorCount);
        }
    AFFT_HIP(hipMalloc(&r.scratch, (size_t)std::max<int64_t>(1, r.scratch_pts) * es));
    return AFFT_OK;
}

static int ragged_launch(afft_ragged_s& r, int dir, void* dst, const void* src, hipStream_t s) {
    void* rargs[] = {&r.counter};
    AFFT_HIP(hipLaunchKernel((const void*)&k_rag_reset, dim3(1), dim3(64), rargs, 0, s));
    for (int p = 0; p < RAG_PHASES; ++p) {
        if (r.ntask[p] == 0) continue;
        RagArgs a{};
        a.in = src;
        a.out = dst;
        a.jobs = r.jobs_dev;
        a.tasks = r.tasks_dev + r.task_off[p];
        a.ntasks = r.ntask[p];
        a.dir = dir;
        a.counter = r.counter + p;
        a.scratch = r.scratch;
        void* args[] = {&a};
        const int grid = std::min(r.grid[dir > 0 ? 1 : 0][p], r.ntask[p]);
        AFFT_HIP(hipLaunchKernel(rag_kernel(r.prec, p, dir), dim3(grid), dim3(RAG_THREADS), args, 0, s));
    }
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
