// Autotuner and wisdom: times the candidate strategies of a shape and keeps
// the winner as a wisdom record, in the text format of the reference
// (internal/planner/wisdom.go:96-205: "size:precision:features:algorithm:timestamp",
// sorted by size, precision, features).  The timing loop replaces the
// reference's MeasureAndSelect (internal/fft/measure.go:89-228): HIP events,
// device-resident buffers, warm-up then timed iterations.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <ctime>
#include <sstream>

#include "afft_internal.hpp"

namespace afft {

static bool is_pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }
static int ilog2i(int64_t n) {
    int l = 0;
    while ((1LL << l) < n) ++l;
    return l;
}

std::vector<Candidate> plan_candidates(const afft_plan_desc& d) {
    std::vector<Candidate> out;
    if (d.kind != AFFT_C2C || d.rank != 1) return out;
    const int64_t n = d.dims[0];
    const int prec = d.precision;
    if (n < 2) return out;
    const bool smooth = !smooth_schedule(n).empty();
    const int maxp = max_single_pass(prec), gcap = gen_cap(prec);
    if (is_pow2(n) && n <= maxp) out.push_back({AFFT_STRATEGY_STOCKHAM, 0});
    const bool rows = d.istride <= 1 && d.ostride <= 1;
    if (smooth && (n <= gcap || (rows && n <= gen_row_cap(prec)))) out.push_back({AFFT_STRATEGY_MIXED, 0});
    if (is_pow2(n)) {
        const int lg = ilog2i(n);
        for (int a = 4; a <= lg - 4; ++a) {
            const int64_t n1 = 1LL << a, n2 = n >> a;
            if (n1 > maxp || n2 > maxp) continue;
            if (std::abs(2 * a - lg) > 4) continue;   // near-balanced splits only
            out.push_back({AFFT_STRATEGY_SIXSTEP, n1});
        }
    } else if (smooth) {
        std::vector<int64_t> divs;
        for (int64_t q = 2; q * q <= n; ++q)
            if (n % q == 0 && q <= gcap && n / q <= gcap) divs.push_back(q);
        // the three most balanced factorisations
        std::sort(divs.begin(), divs.end(), [&](int64_t a, int64_t b) {
            return std::fabs(std::log((double)a * a / n)) < std::fabs(std::log((double)b * b / n));
        });
        for (size_t i = 0; i < divs.size() && i < 3; ++i) out.push_back({AFFT_STRATEGY_SIXSTEP, divs[i]});
        // the splits whose column or row factor is the power-of-two part of
        // n (that pass runs on the compile-time Stockham kernel), and the
        // built-in routing's default split
        auto add = [&](int64_t s) {
            bool seen = false;
            for (const Candidate& c : out) seen |= c.strategy == AFFT_STRATEGY_SIXSTEP && c.split == s;
            if (!seen) out.push_back({AFFT_STRATEGY_SIXSTEP, s});
        };
        const int64_t p2 = n & -n;
        for (int64_t s : {p2, n / p2}) {
            const int64_t o = n / s;
            if (p2 < 16 || s > (s == p2 ? maxp : gcap) || o > (o == p2 ? maxp : gcap)) continue;
            add(s);
        }
        const int64_t s0 = gen_fourstep_split(n, prec);
        if (s0 > 1 && s0 < n) add(s0);
    }
    // Bluestein is the only route for other sizes; for smooth sizes it is a
    // candidate while its padded length stays modest
    if (!smooth || (!is_pow2(n) && 2 * n - 1 <= (1 << 20))) out.push_back({AFFT_STRATEGY_BLUESTEIN, 0});
    return out;
}

// Wisdom feature word of a plan: the device's architecture number in the low
// 16 bits (0x950) and the shape class above it, so that a winner timed at one
// batch size and layout is applied only to plans of the same class: batch
// bucket 0 (1), 1 (2-15), 2 (16-255), 3 (256-4095), 4 (4096 and more), plus 8
// for strided layouts.  The reference's field holds CPU feature bits
// (planner/wisdom.go:17-30); a file written without the class (features
// 0x950) applies to single unit-stride transforms.
uint64_t wisdom_features(const Context& c, const afft_plan_desc& d) {
    const int64_t b = d.batch > 0 ? d.batch : 1;
    const int bucket = b <= 1 ? 0 : b < 16 ? 1 : b < 256 ? 2 : b < 4096 ? 3 : 4;
    const bool strided = d.istride > 1 || d.ostride > 1;
    return c.features() | ((uint64_t)(bucket + (strided ? 8 : 0)) << 16);
}

std::string strategy_name(const Candidate& c) {
    switch (c.strategy) {
    case AFFT_STRATEGY_STOCKHAM: return "hip_stockham";
    case AFFT_STRATEGY_MIXED: return "hip_mixed";
    case AFFT_STRATEGY_BLUESTEIN: return "hip_bluestein";
    case AFFT_STRATEGY_SIXSTEP: return "hip_sixstep_" + std::to_string(c.split);
    }
    return "hip_auto";
}

bool parse_strategy(const std::string& a, Candidate* c) {
    if (a == "hip_stockham") { *c = {AFFT_STRATEGY_STOCKHAM, 0}; return true; }
    if (a == "hip_mixed") { *c = {AFFT_STRATEGY_MIXED, 0}; return true; }
    if (a == "hip_bluestein") { *c = {AFFT_STRATEGY_BLUESTEIN, 0}; return true; }
    if (a.rfind("hip_sixstep_", 0) == 0) {
        *c = {AFFT_STRATEGY_SIXSTEP, std::atoll(a.c_str() + 12)};
        return c->split > 1;
    }
    return false;
}

// 32 features of (shape, candidate); the cost model standardises them.
void plan_features(const afft_plan_desc& d, const Candidate& c, float* f) {
    for (int i = 0; i < NFEAT; ++i) f[i] = 0.f;
    const int64_t n = d.dims[0];
    const int64_t b = d.batch > 0 ? d.batch : 1;
    const int prec = d.precision;
    const double ln = std::log2((double)n), lb = std::log2((double)b);
    int e[4] = {0, 0, 0, 0};
    int64_t m = n;
    const int q[4] = {2, 3, 5, 7};
    for (int i = 0; i < 4; ++i)
        while (m % q[i] == 0) { m /= q[i]; ++e[i]; }
    const bool smooth = m == 1;
    // executed FFT length: the Bluestein padding for a Bluestein candidate
    // (also for smooth n), else n
    const bool blue = c.strategy == AFFT_STRATEGY_BLUESTEIN || (c.strategy == AFFT_STRATEGY_AUTO && !smooth);
    const int64_t M = blue ? [&] { int64_t v = 1; while (v < 2 * n - 1) v <<= 1; return v; }() : n;
    const int64_t stride = d.istride > 1 ? d.istride : 1;
    int64_t n1 = c.split, n2 = c.split ? n / c.split : 0;
    if (c.strategy == AFFT_STRATEGY_SIXSTEP && !n1) { n1 = 1LL << (ilog2i(n) / 2); n2 = n / n1; }
    const double es = prec ? 16.0 : 8.0;
    int launches = 1;
    if (c.strategy == AFFT_STRATEGY_SIXSTEP) launches = 2;
    if (c.strategy == AFFT_STRATEGY_BLUESTEIN) launches = M <= max_single_pass(prec) ? 1 : 4;
    // radix passes of a length and how many of them are odd (3, 5, 7)
    auto passes = [](int64_t len, int* odd) {
        *odd = 0;
        if (len < 2) return 0;
        std::vector<int> r = smooth_schedule(len);
        for (int x : r) *odd += (x & 1);
        return (int)r.size();
    };
    int odd_n = 0, odd1 = 0, odd2 = 0;
    const int pn = passes(M, &odd_n);
    const int p1 = passes(n1, &odd1), p2 = passes(n2, &odd2);
    const bool split = c.strategy == AFFT_STRATEGY_SIXSTEP;
    f[0] = (float)ln;
    f[1] = (float)lb;
    f[2] = (float)(ln + lb);
    f[3] = (float)prec;
    f[4] = is_pow2(n) ? 1.f : 0.f;
    for (int i = 0; i < 4; ++i) f[5 + i] = (float)e[i];
    f[9] = smooth ? 1.f : 0.f;
    f[10] = (float)(split ? p2 : 0);
    f[11] = (float)std::log2((double)M);
    f[12] = stride > 1 ? 1.f : 0.f;
    f[13] = (float)std::log2((double)stride);
    f[14] = c.strategy == AFFT_STRATEGY_STOCKHAM;
    f[15] = c.strategy == AFFT_STRATEGY_SIXSTEP;
    f[16] = c.strategy == AFFT_STRATEGY_BLUESTEIN;
    f[17] = c.strategy == AFFT_STRATEGY_MIXED;
    f[18] = c.strategy == AFFT_STRATEGY_AUTO;
    f[19] = n1 ? (float)std::log2((double)n1) : 0.f;
    f[20] = n2 ? (float)std::log2((double)n2) : 0.f;
    f[21] = n <= max_single_pass(prec) ? 1.f : 0.f;
    f[22] = (n <= gen_cap(prec) || (stride == 1 && n <= gen_row_cap(prec))) ? 1.f : 0.f;
    f[23] = (float)launches;
    f[24] = (float)std::log2(es * 2.0 * n * b * launches);
    f[25] = (split && is_pow2(n1)) ? 1.f : 0.f;
    f[26] = n1 ? (float)std::fabs(std::log2((double)n1) - std::log2((double)n2)) : 0.f;
    f[27] = (split && is_pow2(n2)) ? 1.f : 0.f;
    f[28] = (float)(split ? p1 : pn);
    f[29] = (float)std::log2(std::max(2.0, (double)M * std::log2((double)std::max<int64_t>(M, 2)) * b));
    f[30] = (float)(split ? odd1 + odd2 : odd_n);
    f[31] = 1.f;
}

}  // namespace afft

using namespace afft;

// Times one plan: device-resident buffers, `warmup` untimed executions, then
// three rounds of `iters` executions between two events on a private stream;
// the fastest round's mean counts.  Small transforms take a few microseconds,
// so one slow launch or a clock change would otherwise decide the pick.
static int time_desc(afft_context ctx, const afft_plan_desc& d, int warmup, int iters, double* ms) {
    afft_plan p = nullptr;
    int rc = afft_plan_create_desc(ctx, &d, &p);
    if (rc) return rc;
    int64_t il = 0, ol = 0;
    afft_plan_io_len(p, AFFT_FORWARD, &il, &ol);
    const size_t es = d.precision ? 16 : 8;
    void *in = nullptr, *out = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&] {
        if (in) (void)hipFree(in);
        if (out) (void)hipFree(out);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
        afft_plan_destroy(p);
    };
    hipError_t e;
    if ((e = hipMalloc(&in, il * es)) != hipSuccess || (e = hipMalloc(&out, ol * es)) != hipSuccess ||
        (e = hipMemset(in, 0, il * es)) != hipSuccess || (e = hipStreamCreate(&s)) != hipSuccess ||
        (e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
        cleanup();
        return set_error(AFFT_ERR_DEVICE, std::string("tuner setup: ") + hipGetErrorString(e));
    }
    for (int i = 0; i < warmup && !rc; ++i) rc = afft_plan_exec_device(p, AFFT_FORWARD, out, in, s);
    double best = -1;
    for (int round = 0; round < 3 && !rc; ++round) {
        (void)hipEventRecord(e0, s);
        for (int i = 0; i < iters && !rc; ++i) rc = afft_plan_exec_device(p, AFFT_FORWARD, out, in, s);
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        float f = 0;
        (void)hipEventElapsedTime(&f, e0, e1);
        const double t = f / std::max(1, iters);
        if (best < 0 || t < best) best = t;
    }
    if (!rc) *ms = best;
    cleanup();
    return rc;
}

extern "C" {

int afft_plan_candidates(const afft_plan_desc* d, int* strategies, int64_t* splits, int cap) {
    if (!d) return -1;
    std::vector<Candidate> c = plan_candidates(*d);
    for (int i = 0; i < (int)c.size() && i < cap; ++i) {
        if (strategies) strategies[i] = c[i].strategy;
        if (splits) splits[i] = c[i].split;
    }
    return (int)c.size();
}

int afft_plan_features(const afft_plan_desc* d, int strategy, int64_t split, float* out) {
    if (!d || !out) return set_error(AFFT_ERR_NIL_SLICE, "NULL argument");
    plan_features(*d, Candidate{strategy, split}, out);
    return AFFT_OK;
}

int afft_tune_time(afft_context ctx, const afft_plan_desc* d, int warmup, int iters, double* ms) {
    if (!ctx || !d || !ms) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL argument");
    AFFT_HIP(hipSetDevice(ctx->device));
    int rc = time_desc(ctx, *d, warmup, iters, ms);
    if (!rc) clear_error();
    return rc;
}

// Exhaustive timing of every candidate (MeasureAndSelect); the winner is
// stored as wisdom in the context.  times[i] receives candidate i's time
// (ms, or -1 if it could not be built).
int afft_tune_plan(afft_context ctx, const afft_plan_desc* d, int warmup, int iters, int* best_strategy,
                   int64_t* best_split, double* best_ms, double* times, int cap) {
    if (!ctx || !d) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL argument");
    AFFT_HIP(hipSetDevice(ctx->device));
    std::vector<Candidate> cand = plan_candidates(*d);
    if (cand.empty()) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "no candidates for this shape");
    double best = 1e300;
    Candidate win{0, 0};
    for (size_t i = 0; i < cand.size(); ++i) {
        afft_plan_desc q = *d;
        q.strategy = cand[i].strategy;
        q.split = cand[i].split;
        double t = -1;
        if (time_desc(ctx, q, warmup, iters, &t) != AFFT_OK) t = -1;
        if (times && (int)i < cap) times[i] = t;
        if (t > 0 && t < best) {
            best = t;
            win = cand[i];
        }
    }
    if (best >= 1e299) return set_error(AFFT_ERR_DEVICE, "no candidate could be timed");
    if (best_strategy) *best_strategy = win.strategy;
    if (best_split) *best_split = win.split;
    if (best_ms) *best_ms = best;
    {
        std::lock_guard<std::mutex> l(ctx->mu);
        WisdomEntry w;
        w.size = d->dims[0];
        w.precision = d->precision;
        w.features = wisdom_features(*ctx, *d);
        w.algorithm = strategy_name(win);
        w.timestamp = (int64_t)std::time(nullptr);
        ctx->wisdom[std::make_tuple(w.size, w.precision, w.features)] = w;
    }
    clear_error();
    return AFFT_OK;
}

uint64_t afft_wisdom_features(afft_context ctx, const afft_plan_desc* d) {
    if (!ctx || !d) return 0;
    return wisdom_features(*ctx, *d);
}

// Wisdom.Export: sorted lines "size:precision:features:algorithm:timestamp".
int afft_wisdom_export(afft_context ctx, char* buf, int64_t cap, int64_t* needed) {
    if (!ctx) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL context");
    std::ostringstream os;
    {
        std::lock_guard<std::mutex> l(ctx->mu);
        for (auto& kv : ctx->wisdom) {   // std::map order = (size, precision, features)
            const WisdomEntry& w = kv.second;
            os << w.size << ':' << w.precision << ':' << w.features << ':' << w.algorithm << ':' << w.timestamp << '\n';
        }
    }
    const std::string s = os.str();
    if (needed) *needed = (int64_t)s.size() + 1;
    if (buf && cap > 0) {
        const size_t k = std::min<size_t>((size_t)cap - 1, s.size());
        std::memcpy(buf, s.data(), k);
        buf[k] = 0;
        if ((int64_t)s.size() + 1 > cap) return set_error(AFFT_ERR_LENGTH_MISMATCH, "buffer too small");
    }
    clear_error();
    return AFFT_OK;
}

// Wisdom.Import: blank lines and '#' comments are skipped; a malformed line
// is an error (planner/wisdom.go:149-205).
int afft_wisdom_import(afft_context ctx, const char* text) {
    if (!ctx || !text) return set_error(AFFT_ERR_NIL_SLICE, "NULL argument");
    std::istringstream is(text);
    std::string line;
    std::vector<WisdomEntry> add;
    int lineno = 0;
    while (std::getline(is, line)) {
        ++lineno;
        if (line.empty() || line[0] == '#') continue;
        std::vector<std::string> parts;
        std::string cur;
        std::istringstream ls(line);
        while (std::getline(ls, cur, ':')) parts.push_back(cur);
        if (parts.size() != 5) return set_error(AFFT_ERR_INVALID_ARGUMENT, "wisdom line " + std::to_string(lineno));
        WisdomEntry w;
        char* end = nullptr;
        w.size = std::strtoll(parts[0].c_str(), &end, 10);
        if (*end || w.size < 1) return set_error(AFFT_ERR_INVALID_ARGUMENT, "wisdom size, line " + std::to_string(lineno));
        w.precision = (int)std::strtol(parts[1].c_str(), &end, 10);
        if (*end || (w.precision != 0 && w.precision != 1))
            return set_error(AFFT_ERR_INVALID_ARGUMENT, "wisdom precision, line " + std::to_string(lineno));
        w.features = std::strtoull(parts[2].c_str(), &end, 10);
        if (*end) return set_error(AFFT_ERR_INVALID_ARGUMENT, "wisdom features, line " + std::to_string(lineno));
        w.algorithm = parts[3];
        Candidate c;
        if (!parse_strategy(w.algorithm, &c))
            return set_error(AFFT_ERR_INVALID_ARGUMENT, "wisdom algorithm, line " + std::to_string(lineno));
        w.timestamp = std::strtoll(parts[4].c_str(), &end, 10);
        if (*end) return set_error(AFFT_ERR_INVALID_ARGUMENT, "wisdom timestamp, line " + std::to_string(lineno));
        add.push_back(w);
    }
    std::lock_guard<std::mutex> l(ctx->mu);
    for (auto& w : add) ctx->wisdom[std::make_tuple(w.size, w.precision, w.features)] = w;
    clear_error();
    return AFFT_OK;
}

int afft_wisdom_clear(afft_context ctx) {
    if (!ctx) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL context");
    std::lock_guard<std::mutex> l(ctx->mu);
    ctx->wisdom.clear();
    clear_error();
    return AFFT_OK;
}

}  // extern "C"
