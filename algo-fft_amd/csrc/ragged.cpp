// Ragged batches: many independent 1-D complex transforms of different
// lengths, packed back to back (job i at offset sum of the earlier lengths).
//
// Three ways to run the same batch:
//   afft_ragged_exec        the persistent work-queue kernels, one launch per
//                           task type (column tasks, row tasks, the
//                           Bluestein jobs' second column tasks, whole jobs;
//                           plain and Bluestein jobs apart; power-of-two
//                           sub-transforms per length), after one
//                           counter-reset kernel
//   afft_ragged_exec_graph  the same launches captured once in a hipGraph
//                           and replayed as one graph launch
//   afft_ragged_exec_per_size  one batched plan execution per run of
//                           consecutive jobs of equal length: with the jobs
//                           sorted by length, one launch sequence per
//                           distinct size (the plain per-size path the
//                           persistent kernel is measured against)
//   afft_ragged_exec_per_job   every job through the plan of its own size,
//                           one plan execution per job
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>

#include "afft_internal.hpp"
#include "kragged.hpp"

using namespace afft;

namespace afft {
const void* rag_p2_kernel(int prec, int kind, int L, int dir);   // ragged_p2.hip
const void* rag_whole_kernel(int prec, int kind, int L, int dir, int* threads);   // ragged_whole.hip
int rag_whole_p2_min();
}

namespace afft {
// Clears the launch counters (a kernel, so the launches can be captured in a
// graph and replayed).
__global__ void k_rag_reset(int* counter) {
    if (threadIdx.x < RAG_COUNTERS) counter[threadIdx.x] = 0;
}
}  // namespace afft

// The launches form independent chains (the tasks of a job depend only on
// earlier launches of the same chain): split runtime-pass lengths, power-of-
// two lengths (split, then whole), split Bluestein jobs, and the whole jobs
// of the runtime passes and of Bluestein.  Whole jobs depend on nothing, so
// they can go on any chain; the fourth chain takes the ones that lengthened
// the longest chains.  Measured on the 4096-job mixed set: 0.561 ms with the
// whole Bluestein jobs at the tail of the split Bluestein chain (four
// launches of few blocks each added about 95 us of mostly idle device to
// it), 0.515-0.517 ms with them on a chain of their own, 0.478-0.481 ms
// once the split kernels had sped up, and 0.468-0.470 ms with the runtime-
// pass whole jobs moved there too from the tail of the split runtime-pass
// chain (then the last to finish); 0.604 ms with every whole-job launch on
// a chain of its own.
constexpr int RAG_BRANCHES = 4;
static int rag_branch(int kind) {
    switch (kind) {
    case RL_COLS:
    case RL_ROWS: return 0;
    case RL_COLS_P2:
    case RL_ROWS_P2:
    case RL_WHOLE_P2: return 1;
    case RL_WHOLE:
    case RL_WHOLE_BLUE: return 3;
    default: return 2;   // RL_COLS_BLUE, RL_ROWS_BLUE, RL_COLS2
    }
}

// One launch of the batch: a task type (RL_*), the sub-transform length of
// the power-of-two types (0 for the others), its slice of the task array and
// its persistent grid per direction.
struct RagLaunch {
    int kind = 0, L = 0;
    int ntask = 0, task_off = 0;
    int threads = RAG_THREADS;
    int grid[2] = {0, 0};   // [inverse]
};

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
    int* joblist_dev = nullptr;   // job ids of the grouped whole-job tasks
    int njobs = 0;
    std::vector<RagLaunch> launches;    // in launch order
    int* counter = nullptr;
    void* scratch = nullptr;
    int64_t scratch_pts = 0;
    hipStream_t stream = nullptr;
    // Independent chains of launches run as concurrent branches: the caller's
    // stream takes branch 0, these streams the others; events fork and join.
    hipStream_t branch[RAG_BRANCHES - 1] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[RAG_BRANCHES - 1] = {};
    // Executions share the counters and the scratch array: one enqueued on
    // another stream than the previous one waits for it first.
    hipEvent_t done = nullptr;
    hipStream_t last = nullptr;
    bool ran = false;
    std::map<std::tuple<int, void*, const void*, void*>, hipGraphExec_t> graphs;
    // plain plans of the comparison forms, by (length, batch)
    std::map<std::pair<int64_t, int64_t>, afft_plan> per_job;
    std::map<std::pair<int64_t, int64_t>, afft_plan> per_size;
    std::mutex mu;
    ~afft_ragged_s();
};

afft_ragged_s::~afft_ragged_s() {
    if (ctx) (void)hipSetDevice(ctx->device);
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    for (auto& kv : per_job) afft_plan_destroy(kv.second);
    for (auto& kv : per_size) afft_plan_destroy(kv.second);
    if (jobs_dev) (void)hipFree(jobs_dev);
    if (tasks_dev) (void)hipFree(tasks_dev);
    if (joblist_dev) (void)hipFree(joblist_dev);
    if (counter) (void)hipFree(counter);
    if (scratch) (void)hipFree(scratch);
    if (stream) (void)hipStreamDestroy(stream);
    for (int b = 0; b < RAG_BRANCHES - 1; ++b) {
        if (branch[b]) (void)hipStreamDestroy(branch[b]);
        if (join[b]) (void)hipEventDestroy(join[b]);
    }
    if (fork) (void)hipEventDestroy(fork);
    if (done) (void)hipEventDestroy(done);
}

// Orders an execution on s after the previous one (outside stream capture,
// where the caller orders the graph's nodes).
static int ragged_order_begin(afft_ragged_s& r, hipStream_t s, bool* track) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    AFFT_HIP(hipStreamIsCapturing(s, &cs));
    *track = cs == hipStreamCaptureStatusNone;
    if (*track && r.ran && r.last != s) AFFT_HIP(hipStreamWaitEvent(s, r.done, 0));
    return AFFT_OK;
}

static int ragged_order_end(afft_ragged_s& r, hipStream_t s, bool track) {
    if (!track) return AFFT_OK;
    AFFT_HIP(hipEventRecord(r.done, s));
    r.last = s;
    r.ran = true;
    return AFFT_OK;
}

// Radix schedule and per-pass twiddles of length m (m <= rag_whole_cap), shared
// between jobs through `cache`.  The table is r-major (entry (r-1)*Ns + k,
// the layout of the compile-time cores) for the power-of-two split tasks and
// k-major (entry k*(R-1) + r-1, one base address per butterfly) for the
// runtime passes.
using SchedKey = std::pair<int64_t, bool>;
static int schedule_into(Plan& tab, int prec, int64_t m, bool rmajor, RSched& j, std::map<SchedKey, RSched>& cache) {
    const SchedKey key{m, rmajor};
    auto it = cache.find(key);
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
        cache[key] = j;
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
            const int R = rad[i];
            for (int64_t e = 0; e < ns * (R - 1); ++e) {
                const int64_t k = rmajor ? e % ns : e / (R - 1);
                const int r = rmajor ? (int)(e / ns) + 1 : (int)(e % (R - 1)) + 1;
                double re, im;
                root_of_unity(k * r, ns * R, &re, &im);
                tw.push_back(re);
                tw.push_back(im);
            }
            off += (int)(ns * (R - 1));
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
    cache[key] = j;
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

// The persistent kernel of a launch for direction dir.
template <typename T>
static const void* rag_kernel_t(int kind, int dir) {
    switch (kind) {
    case RL_COLS: return dir > 0 ? (const void*)&k_rag_cols<T, 1> : (const void*)&k_rag_cols<T, -1>;
    case RL_ROWS: return dir > 0 ? (const void*)&k_rag_rows<T, 1> : (const void*)&k_rag_rows<T, -1>;
    default: return nullptr;
    }
}

static bool whole_kind(int kind) { return kind == RL_WHOLE_P2 || kind == RL_WHOLE || kind == RL_WHOLE_BLUE; }

static const void* rag_kernel(int prec, const RagLaunch& l, int dir) {
    int th = 0;
    if (whole_kind(l.kind)) return rag_whole_kernel(prec, l.kind, l.L, dir, &th);
    if (l.L) return rag_p2_kernel(prec, l.kind, l.L, dir);
    return prec ? rag_kernel_t<double>(l.kind, dir) : rag_kernel_t<float>(l.kind, dir);
}

static int rag_threads(int prec, const RagLaunch& l) {
    int th = RAG_THREADS;
    if (whole_kind(l.kind)) (void)rag_whole_kernel(prec, l.kind, l.L, -1, &th);
    else if (l.L) th = prec ? RagS<double>::THREADS : RagS<float>::THREADS;   // power-of-two split tasks
    else if (l.kind == RL_COLS || l.kind == RL_ROWS) th = prec ? RagG<double>::THREADS : RagG<float>::THREADS;
    return th;
}

// A split runs on the power-of-two cores when both lengths are powers of two
// in the ranges those kernels are built for, with tasks of pts/L
// sub-transforms of exactly pts points (RagS<T>::PTS).
static bool p2_split(int64_t m1, int64_t m2, int pts) {
    auto pow2 = [](int64_t L) { return L > 0 && (L & (L - 1)) == 0; };
    return pow2(m1) && pow2(m2) && m1 >= RAG_P2_COL_MIN && m1 <= RAG_P2_COL_MAX && m2 >= RAG_P2_ROW_MIN &&
           m2 <= RAG_P2_ROW_MAX && m2 % (pts / m1) == 0 && m1 % (pts / m2) == 0;
}

static int p2_pts(int prec) { return prec ? RagS<double>::PTS : RagS<float>::PTS; }


static int ragged_build(afft_ragged_s& r) {
    const int es = r.prec ? 16 : 8;
    std::vector<RJob> jobs(r.sizes.size());
    std::map<SchedKey, RSched> scache;
    std::map<int, std::vector<int>> groups;       // whole power-of-two jobs by length
    std::map<int, std::vector<int>> gen_groups;   // other whole plain jobs by length
    std::map<int, std::vector<int>> blue_groups;  // whole Bluestein jobs by padded length
    const int wpts = r.prec ? RagW<double>::PTS : RagW<float>::PTS;   // points per whole-job task
    const int wth = r.prec ? RagW<double>::THREADS : RagW<float>::THREADS;
    std::map<std::pair<int64_t, int64_t>, BlueTables> bcache;
    struct TwSplit { int shift; const void* lo; const void* hi; };
    std::map<int64_t, TwSplit> tcache;
    struct Pending { RTask t; double cost; };
    // tasks per launch, keyed by (task type, sub-transform length): the
    // map's order is the launch order
    std::map<std::pair<int, int>, std::vector<Pending>> ph;
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
        int64_t cap = rag_whole_cap(r.prec);
        bool blue = !(n == 1 || !smooth_schedule(n).empty());
        if (!blue && m > cap && !split_len(m, &m1, &m2)) blue = true;   // no usable split
        if (blue) {
            // Bluestein paddings above 8192 points split: a whole one holds
            // two transforms' live values and spills at 1024 threads
            // (measured 0.555 against 0.560 ms on the 4096-job set)
            cap = std::min<int64_t>(cap, 8192);
            m = next_pow2(2 * n - 1);
            if (m > cap) split_len(m, &m1, &m2);
        }
        j.m = (int)m;
        j.blue = blue;
        const bool whole = m <= cap;
        int rc;
        // whole power-of-two jobs and Bluestein paddings of rag_whole_p2_min()
        // points or more run on the compile-time cores (r-major table), the
        // rest on the runtime passes
        const bool whole_p2 = whole && m >= rag_whole_p2_min() && (m & (m - 1)) == 0;
        if (whole) {
            if ((rc = schedule_into(r.tables, r.prec, m, whole_p2, j.s[0], scache))) return rc;
        } else {
            j.m1 = (int)m1;
            j.m2 = (int)m2;
            // power-of-two splits: p2_pts points per task on the compile-time
            // cores; generic splits: at most RagG<T>::PTS points per task, so that
            // every thread holds at most 16/R butterflies of a radix-R pass
            const bool rmaj = p2_split(m1, m2, p2_pts(r.prec));
            if (rmaj) {
                j.g_col = p2_pts(r.prec) / (int)m1;
                j.g_row = p2_pts(r.prec) / (int)m2;
            } else {
                const int gth = r.prec ? RagG<double>::THREADS : RagG<float>::THREADS;
                const int gpts = r.prec ? RagG<double>::PTS : RagG<float>::PTS;
                j.g_col = pow2_floor(std::min<int64_t>({gth, gpts / m1, m2}));
                j.g_row = pow2_floor(std::min<int64_t>({gth, gpts / m2, m1}));
            }
            if ((rc = schedule_into(r.tables, r.prec, m1, rmaj, j.s[0], scache))) return rc;
            if ((rc = schedule_into(r.tables, r.prec, m2, rmaj, j.s[1], scache))) return rc;
            auto tt = tcache.find(m);
            if (tt == tcache.end()) {
                TwSplit ts{};
                if ((rc = split_twiddles(r.tables, r.prec, m, &ts.shift, &ts.lo, &ts.hi))) return rc;
                tt = tcache.emplace(m, ts).first;
            }
            j.shift = tt->second.shift;
            j.twlo = tt->second.lo;
            j.twhi = tt->second.hi;
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
        if (whole_p2) {
            (blue ? blue_groups : groups)[(int)m].push_back(job);
            continue;
        }
        if (whole && blue) {
            // one job per block on the runtime passes
            ph[{RL_WHOLE_BLUE, 0}].push_back({{job, RT_WHOLE, 0, 1}, 2.0 * m * std::log2((double)m)});
            continue;
        }
        if (whole) {
            // wpts points per task at most, TPT = wth/G threads a job (at
            // most 16 points per thread)
            j.g_col = pow2_floor(std::min<int64_t>(wth, wpts / m));
            gen_groups[(int)m].push_back(job);
            continue;
        }
        // Power-of-two splits (every Bluestein job) run on the compile-time
        // cores, which take exactly p2_pts/L sub-transforms per task.
        const bool p2 = p2_split(m1, m2, p2_pts(r.prec));
        if (blue && !p2) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "no power-of-two split for Bluestein length");
        const int lc = blue ? RL_COLS_BLUE : p2 ? RL_COLS_P2 : RL_COLS;
        const int lr = blue ? RL_ROWS_BLUE : p2 ? RL_ROWS_P2 : RL_ROWS;
        // the generic passes stage through LDS: about one pass more traffic
        const double extra = p2 ? 1.0 : 2.0;
        const double ccol = (double)m1 * (std::log2((double)m1) + extra);
        const double crow = (double)m2 * (std::log2((double)m2) + extra) * (blue ? 2.0 : 1.0);
        const int l1 = p2 ? (int)m1 : 0, l2 = p2 ? (int)m2 : 0;
        for (int64_t c0 = 0; c0 < m2; c0 += j.g_col) {
            const int cnt = (int)std::min<int64_t>(j.g_col, m2 - c0);
            ph[{lc, l1}].push_back({{job, RT_COLS, (int)c0, cnt}, ccol * cnt});
            if (blue) ph[{RL_COLS2, l1}].push_back({{job, RT_COLS2, (int)c0, cnt}, ccol * cnt});
        }
        for (int64_t r0 = 0; r0 < m1; r0 += j.g_row) {
            const int cnt = (int)std::min<int64_t>(j.g_row, m1 - r0);
            ph[{lr, l2}].push_back({{job, RT_ROWS, (int)r0, cnt}, crow * cnt});
        }
    }
    r.total = off;
    // grouped whole jobs, ids in joblist: max(1, RAG_PTS/L) jobs of length L
    // per task on the compile-time cores (one launch per length), g_col on
    // the runtime passes (one launch for every length)
    std::vector<int> joblist;
    for (int kind : {RL_WHOLE_P2, RL_WHOLE_BLUE, RL_WHOLE}) {
        for (auto& g : kind == RL_WHOLE_P2 ? groups : kind == RL_WHOLE_BLUE ? blue_groups : gen_groups) {
            const int L = g.first;
            const int per = kind == RL_WHOLE ? jobs[g.second[0]].g_col : std::max(1, RAG_PTS / L);
            const double c = (kind == RL_WHOLE_BLUE ? 2.0 : 1.0) * L * std::log2(std::max(2.0, (double)L));
            for (size_t q = 0; q < g.second.size(); q += per) {
                const int cnt = (int)std::min<size_t>(per, g.second.size() - q);
                const int first = (int)joblist.size();
                joblist.insert(joblist.end(), g.second.begin() + q, g.second.begin() + q + cnt);
                ph[{kind, kind == RL_WHOLE ? 0 : L}].push_back({{g.second[q], RT_WHOLE, first, cnt}, c * cnt});
            }
        }
    }
    // longest-processing-time order within each phase
    std::vector<RTask> tasks;
    r.launches.clear();
    for (auto& kv : ph) {
        auto& q = kv.second;
        std::stable_sort(q.begin(), q.end(), [](const Pending& a, const Pending& b) { return a.cost > b.cost; });
        RagLaunch l;
        l.kind = kv.first.first;
        l.L = kv.first.second;
        l.task_off = (int)tasks.size();
        l.ntask = (int)q.size();
        for (auto& t : q) tasks.push_back(t.t);
        r.launches.push_back(l);
    }
    if ((int)r.launches.size() > RAG_COUNTERS) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "too many ragged launches");
    r.njobs = (int)jobs.size();
    AFFT_HIP(hipMalloc(&r.jobs_dev, std::max<size_t>(1, jobs.size()) * sizeof(RJob)));
    AFFT_HIP(hipMemcpy(r.jobs_dev, jobs.data(), jobs.size() * sizeof(RJob), hipMemcpyHostToDevice));
    AFFT_HIP(hipMalloc(&r.joblist_dev, std::max<size_t>(1, joblist.size()) * sizeof(int)));
    if (!joblist.empty())
        AFFT_HIP(hipMemcpy(r.joblist_dev, joblist.data(), joblist.size() * sizeof(int), hipMemcpyHostToDevice));
    AFFT_HIP(hipMalloc(&r.tasks_dev, std::max<size_t>(1, tasks.size()) * sizeof(RTask)));
    if (!tasks.empty())
        AFFT_HIP(hipMemcpy(r.tasks_dev, tasks.data(), tasks.size() * sizeof(RTask), hipMemcpyHostToDevice));
    AFFT_HIP(hipMalloc(&r.counter, RAG_COUNTERS * sizeof(int)));
    // Each persistent launch gets a share of the device's resident blocks,
    // so that the concurrent branches run side by side instead of the first
    // one filling every CU until its queue drains.  Weights 4:5:7:3 (split
    // runtime passes, power-of-two cores, split Bluestein, whole runtime-pass
    // and Bluestein jobs) over the branches
    // present, 1.7 device-fulls in all, at most one device-full per branch;
    // a set with one branch keeps the whole device.  Shares swept on the
    // 4096-job mixed set: 0.561-0.61 ms over the settings tried with three
    // chains; with the current chains a fourth-chain weight of 2 to 4 gives
    // 0.468-0.469 ms, 5 or 1.9 device-fulls 0.473 ms.
    static constexpr int weight[RAG_BRANCHES] = {4, 5, 7, 3};
    bool present[RAG_BRANCHES] = {};
    for (const RagLaunch& l : r.launches) present[rag_branch(l.kind)] = true;
    int wsum = 0;
    for (int b = 0; b < RAG_BRANCHES; ++b) wsum += present[b] ? weight[b] : 0;
    for (int inv = 0; inv < 2; ++inv)
        for (RagLaunch& l : r.launches) {
            const void* fn = rag_kernel(r.prec, l, inv ? 1 : -1);
            if (!fn) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "no ragged kernel for a task type");
            l.threads = rag_threads(r.prec, l);
            int per_cu = 0;
            AFFT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, l.threads, 0));
            const int full = std::max(1, per_cu) * std::max(1, r.ctx->props.multiProcessorCount);
            const double share = std::min(1.0, 1.7 * weight[rag_branch(l.kind)] / std::max(1, wsum));
            l.grid[inv] = std::max(1, (int)(full * share));
        }
    AFFT_HIP(hipMalloc(&r.scratch, (size_t)std::max<int64_t>(1, r.scratch_pts) * es));
    return AFFT_OK;
}

// Enqueues the batch on s: the counter reset, then each chain of launches on
// its own branch (branch 0 on s itself), joined back into s.  Works under
// stream capture as well, where the events become graph edges.
static int ragged_launch(afft_ragged_s& r, int dir, void* dst, const void* src, hipStream_t s) {
    void* rargs[] = {&r.counter};
    AFFT_HIP(hipLaunchKernel((const void*)&k_rag_reset, dim3(1), dim3(RAG_COUNTERS), rargs, 0, s));
    bool used[RAG_BRANCHES] = {};
    for (const RagLaunch& l : r.launches) used[rag_branch(l.kind)] = true;
    AFFT_HIP(hipEventRecord(r.fork, s));
    for (int b = 1; b < RAG_BRANCHES; ++b)
        if (used[b]) AFFT_HIP(hipStreamWaitEvent(r.branch[b - 1], r.fork, 0));
    // Enqueue the split runtime-pass chain first (its generic kernels run at
    // about half the per-point rate of the compile-time ones), then split
    // Bluestein, power-of-two, whole jobs; each branch keeps its own launch
    // order.  Measured on the 4096-job mixed set: 0.508 ms against
    // 0.515-0.518 ms with split Bluestein first; shares of 1.5 to 3.0
    // device-fulls 0.507-0.548 ms.
    static constexpr int order[RAG_BRANCHES] = {0, 2, 1, 3};
    for (int ob = 0; ob < RAG_BRANCHES; ++ob)
    for (size_t p = 0; p < r.launches.size(); ++p) {
        const RagLaunch& l = r.launches[p];
        const int b = rag_branch(l.kind);
        if (b != order[ob]) continue;
        hipStream_t bs = b ? r.branch[b - 1] : s;
        RagArgs a{};
        a.in = src;
        a.out = dst;
        a.jobs = r.jobs_dev;
        a.tasks = r.tasks_dev + l.task_off;
        a.joblist = r.joblist_dev;
        a.ntasks = l.ntask;
        a.dir = dir;
        a.counter = r.counter + p;
        a.scratch = r.scratch;
        void* args[] = {&a};
        const int grid = std::min(l.grid[dir > 0 ? 1 : 0], l.ntask);
        AFFT_HIP(hipLaunchKernel(rag_kernel(r.prec, l, dir), dim3(grid), dim3(l.threads), args, 0, bs));
    }
    for (int b = 1; b < RAG_BRANCHES; ++b)
        if (used[b]) {
            AFFT_HIP(hipEventRecord(r.join[b - 1], r.branch[b - 1]));
            AFFT_HIP(hipStreamWaitEvent(s, r.join[b - 1], 0));
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
        for (int b = 0; b < RAG_BRANCHES - 1 && e == hipSuccess; ++b) {
            e = hipStreamCreateWithFlags(&r->branch[b], hipStreamNonBlocking);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&r->join[b], hipEventDisableTiming);
        }
        if (e == hipSuccess) e = hipEventCreateWithFlags(&r->fork, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&r->done, hipEventDisableTiming);
        if (e != hipSuccess) rc = set_error(AFFT_ERR_DEVICE, std::string("ragged streams: ") + hipGetErrorString(e));
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
    hipStream_t s = stream ? (hipStream_t)stream : r->stream;
    bool track = false;
    int rc = ragged_order_begin(*r, s, &track);
    if (rc == AFFT_OK) rc = ragged_launch(*r, direction, dst, src, s);
    if (rc == AFFT_OK) rc = ragged_order_end(*r, s, track);
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
    bool track = false;
    int rc = ragged_order_begin(*r, s, &track);
    if (rc) return rc;
    AFFT_HIP(hipGraphLaunch(it->second, s));
    if ((rc = ragged_order_end(*r, s, track))) return rc;
    clear_error();
    return AFFT_OK;
}

// A plain C2C plan of length n over `batch` contiguous transforms, cached.
static int ragged_plain_plan(afft_ragged_s& r, std::map<std::pair<int64_t, int64_t>, afft_plan>& cache, int64_t n,
                             int64_t batch, afft_plan* out) {
    auto it = cache.find({n, batch});
    if (it == cache.end()) {
        afft_plan_desc d{};
        d.kind = AFFT_C2C;
        d.precision = r.prec;
        d.rank = 1;
        d.dims[0] = n;
        d.batch = batch;
        d.flags = r.flags;
        afft_plan p = nullptr;
        int rc = afft_plan_create_desc(r.ctx, &d, &p);
        if (rc) return rc;
        it = cache.emplace(std::make_pair(n, batch), p).first;
    }
    *out = it->second;
    return AFFT_OK;
}

int afft_ragged_exec_per_size(afft_ragged r, int direction, void* dst, const void* src, void* stream) {
    if (!r) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL handle");
    if (!dst || !src) return set_error(AFFT_ERR_NIL_SLICE, "dst or src is nil");
    if (direction != AFFT_FORWARD && direction != AFFT_INVERSE)
        return set_error(AFFT_ERR_INVALID_ARGUMENT, "direction must be -1 or +1");
    std::lock_guard<std::mutex> l(r->mu);
    AFFT_HIP(hipSetDevice(r->ctx->device));
    const int es = r->prec ? 16 : 8;
    const size_t nj = r->sizes.size();
    for (size_t i = 0; i < nj;) {
        // the run of consecutive jobs of this length is one contiguous batch
        size_t e = i + 1;
        while (e < nj && r->sizes[e] == r->sizes[i]) ++e;
        afft_plan p = nullptr;
        int rc = ragged_plain_plan(*r, r->per_size, r->sizes[i], (int64_t)(e - i), &p);
        if (rc) return rc;
        const int64_t o = r->offs[i] * es;
        rc = afft_plan_exec_device(p, direction, (char*)dst + o, (const char*)src + o,
                                   stream ? stream : (void*)r->stream);
        if (rc) return rc;
        i = e;
    }
    clear_error();
    return AFFT_OK;
}

int afft_ragged_exec_per_job(afft_ragged r, int direction, void* dst, const void* src, void* stream) {
    if (!r) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL handle");
    if (!dst || !src) return set_error(AFFT_ERR_NIL_SLICE, "dst or src is nil");
    if (direction != AFFT_FORWARD && direction != AFFT_INVERSE)
        return set_error(AFFT_ERR_INVALID_ARGUMENT, "direction must be -1 or +1");
    std::lock_guard<std::mutex> l(r->mu);
    AFFT_HIP(hipSetDevice(r->ctx->device));
    const int es = r->prec ? 16 : 8;
    for (size_t i = 0; i < r->sizes.size(); ++i) {
        const int64_t n = r->sizes[i];
        afft_plan p = nullptr;
        int rc = ragged_plain_plan(*r, r->per_job, n, 1, &p);
        if (rc) return rc;
        const int64_t o = r->offs[i] * es;
        rc = afft_plan_exec_device(p, direction, (char*)dst + o, (const char*)src + o,
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
