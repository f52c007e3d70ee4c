// Planner and executor: turns a plan descriptor into a list of kernel
// launches (steps) and runs them.
//
// Size routing (cf. the reference's planner.go:43-78, which picks codelets,
// Stockham, six-step or Bluestein for a CPU):
//   power of two <= LDS capacity (16384 c64 / 8192 c128)  one LDS pass
//   power of two above it                                 four-step: column pass with
//                                                         fused twiddle, then a row pass
//                                                         whose store transposes
//   2/3/5/7-smooth                                        runtime mixed-radix Stockham
//   anything else                                         fused Bluestein chain
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include <cmath>

#include "afft_internal.hpp"
#include "kgen.hpp"

namespace afft {

Plan::~Plan() {
    if (ctx) (void)hipSetDevice(ctx->device);
    for (void* q : owned) (void)hipFree(q);
    for (void*& t : tmp)
        if (t) (void)hipFree(t);
    if (stage_in) (void)hipFree(stage_in);
    if (stage_out) (void)hipFree(stage_out);
    if (full) (void)hipFree(full);
    if (pipe) destroy_host_pipe(pipe);
    if (ws_done) (void)hipEventDestroy(ws_done);
    if (stream) (void)hipStreamDestroy(stream);
}

static bool is_pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }
static int ilog2(int64_t n) {
    int l = 0;
    while ((1LL << l) < n) ++l;
    return l;
}

// Upload a host table, owned by the plan.
int upload_table(Plan& p, const void* host, size_t bytes, const void** dev) {
    void* d = nullptr;
    AFFT_HIP(hipMalloc(&d, bytes ? bytes : 16));
    p.owned.push_back(d);
    if (bytes) AFFT_HIP(hipMemcpy(d, host, bytes, hipMemcpyHostToDevice));
    *dev = d;
    return AFFT_OK;
}

// Twiddle table of a radix schedule: for pass i >= 1 with
// Ns = rad[0]*...*rad[i-1] and radix R: entries W_{Ns*R}^{k*r}, k < Ns,
// 1 <= r < R, stored r-major (entry (r-1)*Ns + k), so the lanes of a pass,
// which hold consecutive k, load consecutive entries.  Generated in FP64 and
// rounded once.
static int schedule_twiddles(Plan& p, const KernelInfo* k, const void** dev) {
    std::vector<double> tw;
    int64_t ns = k->rad[0];
    for (int i = 1; i < k->nrad; ++i) {
        const int R = k->rad[i];
        for (int r = 1; r < R; ++r)
            for (int64_t kk = 0; kk < ns; ++kk) {
                double re, im;
                root_of_unity(kk * r, ns * R, &re, &im);
                tw.push_back(re);
                tw.push_back(im);
            }
        ns *= R;
    }
    if (k->prec == 0) {
        std::vector<float> f(tw.begin(), tw.end());
        return upload_table(p, f.data(), f.size() * sizeof(float), dev);
    }
    return upload_table(p, tw.data(), tw.size() * sizeof(double), dev);
}

// Two-level table for W_L^m = hi[m >> sh] * lo[m & (2^sh - 1)].
int split_twiddles(Plan& p, int prec, int64_t L, int* shift, const void** lo, const void** hi) {
    const int sh = (ilog2(L) + 1) / 2;
    const int64_t nlo = 1LL << sh;
    const int64_t nhi = (L + nlo - 1) / nlo;
    std::vector<double> a(2 * nlo), b(2 * nhi);
    for (int64_t i = 0; i < nlo; ++i) root_of_unity(i, L, &a[2 * i], &a[2 * i + 1]);
    for (int64_t i = 0; i < nhi; ++i) root_of_unity(i << sh, L, &b[2 * i], &b[2 * i + 1]);
    *shift = sh;
    int rc;
    if (prec == 0) {
        std::vector<float> fa(a.begin(), a.end()), fb(b.begin(), b.end());
        if ((rc = upload_table(p, fa.data(), fa.size() * 4, lo))) return rc;
        return upload_table(p, fb.data(), fb.size() * 4, hi);
    }
    if ((rc = upload_table(p, a.data(), a.size() * 8, lo))) return rc;
    return upload_table(p, b.data(), b.size() * 8, hi);
}

static Step make_step(const KernelInfo* k, const KArgs& a, int64_t gx, int64_t gy, const char* what) {
    Step s;
    s.fn = k->fn;
    s.grid = dim3((unsigned)gx, (unsigned)gy, 1);
    s.block = dim3((unsigned)k->threads, 1, 1);
    s.arg.resize(sizeof(KArgs));
    std::memcpy(s.arg.data(), &a, sizeof(KArgs));
    s.in_field = (int)offsetof(KArgs, in);
    s.out_field = (int)offsetof(KArgs, out);
    s.nt_field = (int)offsetof(KArgs, nt);
    s.name = std::string(what) + "_n" + std::to_string(k->n) + "_e" + std::to_string(k->e) + "_tpb" +
             std::to_string(k->tpb);
    return s;
}

// Launches with more than 65535 slices along y are split; each piece gets its
// own byte offsets.
static void push_sliced(std::vector<Step>& out, Step s, int64_t slices, int64_t in_bd_bytes, int64_t out_bd_bytes) {
    const int64_t maxy = 65535;
    for (int64_t y0 = 0; y0 < slices; y0 += maxy) {
        Step t = s;
        t.grid.y = (unsigned)std::min(maxy, slices - y0);
        t.in_off = s.in_off + y0 * in_bd_bytes;
        t.out_off = s.out_off + y0 * out_bd_bytes;
        out.push_back(t);
    }
}

// One batched 1-D complex transform of length n along a strided axis:
//   element j of transform (y, c) at  y*bd + c*cd + j*nd   (elements)
// c in [0, count), y in [0, slices).  Emits the steps for direction di.
struct AxisJob {
    int64_t n;
    int64_t count;
    int64_t slices;
    int64_t in_cd, in_nd, in_bd;
    int64_t out_cd, out_nd, out_bd;
    int in_role, out_role;
    int64_t in_off, out_off;   // bytes
    double scale;
    int strategy = AFFT_STRATEGY_AUTO;   // forced route (1-D plans)
    int64_t split = 0;                   // forced four-step n1
};

static int emit_axis(Plan& p, int di, const AxisJob& j, std::string& algo);


// MFMA form (AFFT_FLAG_MFMA): complex64 rows of N = 256 .. 16384 (mfma.hip).
const void* mfma_kernel(int n, int var, int dir, int* per_block, int* threads);

// MFMA form of a single-pass row transform (n = 256 .. 16384).
static int emit_mfma(Plan& p, int di, const AxisJob& j, std::string& algo) {
    const int dir = di ? 1 : -1;
    int g = 1, threads = 256;
    const void* fn = mfma_kernel((int)j.n, VAR_ROW, dir, &g, &threads);
    const KernelInfo* k = find_kernel(0, (int)j.n, VAR_ROW, dir);
    if (p.desc.precision != 0 || !fn || !k || j.in_nd != 1 || j.out_nd != 1)
        return set_error(AFFT_ERR_NOT_IMPLEMENTED,
                         "the MFMA form covers complex64 rows of 2^8 .. 2^24 points");
    KArgs a{};
    int rc = schedule_twiddles(p, k, &a.tw);
    if (rc) return rc;
    a.count = j.count;
    a.in_cd = j.in_cd;
    a.out_cd = j.out_cd;
    a.in_nd = a.out_nd = 1;
    a.scale = j.scale;
    Step s;
    s.fn = fn;
    s.grid = dim3((unsigned)((j.count + g - 1) / g), 1, 1);
    s.block = dim3((unsigned)threads, 1, 1);
    s.arg.resize(sizeof(KArgs));
    std::memcpy(s.arg.data(), &a, sizeof(KArgs));
    s.in_field = (int)offsetof(KArgs, in);
    s.out_field = (int)offsetof(KArgs, out);
    s.nt_field = (int)offsetof(KArgs, nt);
    s.in_role = j.in_role;
    s.out_role = j.out_role;
    s.in_off = j.in_off;
    s.out_off = j.out_off;
    s.name = "mfma" + std::to_string(j.n);
    const int es = 8;
    for (int64_t y = 0; y < j.slices; ++y) {
        Step t = s;
        t.in_off += y * j.in_bd * es;
        t.out_off += y * j.out_bd * es;
        p.steps[di].push_back(t);
    }
    if (!algo.empty()) algo += "+";
    algo += "mfma" + std::to_string(j.n) + "r";
    return AFFT_OK;
}

// A strided axis whose columns are contiguous (cd = 1) and too long for 8
// adjacent columns per block: the column loads would fall apart into
// segments of a few bytes.  Four-step on the axis instead, n = n1*n2, both
// passes on many adjacent columns:
//   A: for each (c, k2): FFT over k1 of x[c + (n2*k1 + k2)*S], times
//      W_n^{k2*j1}, into T[(j1*n2 + k2)*count + c]  (blockIdx.y = k2)
//   B: for each (c, j1): FFT over k2 of T, into X[c + (j1 + n1*j2)*S]
static int emit_col_fourstep(Plan& p, int di, const AxisJob& j, std::string& algo) {
    const int prec = p.desc.precision;
    const int dir = di ? 1 : -1;
    const int es = prec ? 16 : 8;
    const int lg = ilog2(j.n);
    const int64_t n1 = 1LL << (lg / 2), n2 = j.n / n1;
    const KernelInfo* ka = find_kernel(prec, (int)n1, VAR_4A, dir);
    const KernelInfo* kb = find_kernel(prec, (int)n2, VAR_COL, dir);
    if (!ka || !kb) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "missing column four-step kernel");
    const size_t need = (size_t)(j.n * j.count * es);
    if (need > p.tmp_bytes[0]) p.tmp_bytes[0] = need;
    int rc;
    KArgs a{};
    if ((rc = schedule_twiddles(p, ka, &a.tw))) return rc;
    if ((rc = split_twiddles(p, prec, j.n, &a.tw4_shift, &a.tw4_lo, &a.tw4_hi))) return rc;
    a.tw4_L = j.n;
    a.tw4_by_y = 1;
    a.count = j.count;
    a.in_cd = 1;
    a.in_nd = n2 * j.in_nd;
    a.in_bd = j.in_nd;
    a.out_cd = 1;
    a.out_nd = n2 * j.count;
    a.out_bd = j.count;
    a.scale = 1.0;
    KArgs b{};
    if ((rc = schedule_twiddles(p, kb, &b.tw))) return rc;
    b.count = j.count;
    b.in_cd = 1;
    b.in_nd = j.count;
    b.in_bd = n2 * j.count;
    b.out_cd = 1;
    b.out_nd = n1 * j.out_nd;
    b.out_bd = j.out_nd;
    b.scale = j.scale;
    for (int64_t y = 0; y < j.slices; ++y) {
        Step sa = make_step(ka, a, (j.count + ka->tpb - 1) / ka->tpb, n2, "colA");
        sa.in_role = j.in_role;
        sa.in_off = j.in_off + y * j.in_bd * es;
        sa.out_role = ROLE_TMP0;
        sa.out_off = 0;
        p.steps[di].push_back(sa);
        Step sb = make_step(kb, b, (j.count + kb->tpb - 1) / kb->tpb, n1, "colB");
        sb.in_role = ROLE_TMP0;
        sb.in_off = 0;
        sb.out_role = j.out_role;
        sb.out_off = j.out_off + y * j.out_bd * es;
        p.steps[di].push_back(sb);
    }
    if (!algo.empty()) algo += "+";
    algo += "colfourstep" + std::to_string(n1) + "x" + std::to_string(n2);
    return AFFT_OK;
}

static int emit_single(Plan& p, int di, const AxisJob& j, std::string& algo) {
    const int prec = p.desc.precision;
    const int dir = di ? 1 : -1;
    const int es = prec ? 16 : 8;
    if (p.desc.flags & AFFT_FLAG_MFMA) return emit_mfma(p, di, j, algo);
    int var;
    int64_t count = j.count, slices = j.slices;
    int64_t in_cd = j.in_cd, out_cd = j.out_cd, in_bd = j.in_bd, out_bd = j.out_bd;
    if (j.in_nd == 1 && j.out_nd == 1) {
        var = VAR_ROW;
    } else if (j.in_cd == 1 && j.out_cd == 1) {
        var = VAR_COL;
    } else {
        // no unit direction on either axis: one transform per slice
        var = VAR_COL;
        if (count != 1) {
            // fold c into the slice dimension when the slice dimension is trivial
            if (slices != 1) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "layout with no unit stride and two batch axes");
            slices = count;
            in_bd = j.in_cd;
            out_bd = j.out_cd;
            count = 1;
        }
    }
    const KernelInfo* k = find_kernel(prec, (int)j.n, var, dir);
    if (!k) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "no kernel for n=" + std::to_string(j.n));
    // fewer than 8 adjacent columns per block: column loads of under 64 bytes
    if (var == VAR_COL && count >= 16 && k->tpb < 8 && j.n >= 256 && j.in_cd == 1 && j.out_cd == 1 &&
        j.in_nd >= count && j.out_nd >= count)
        return emit_col_fourstep(p, di, j, algo);
    KArgs a{};
    int rc = schedule_twiddles(p, k, &a.tw);
    if (rc) return rc;
    a.count = count;
    a.in_cd = in_cd;
    a.out_cd = out_cd;
    a.in_nd = j.in_nd;
    a.out_nd = j.out_nd;
    a.in_bd = in_bd;
    a.out_bd = out_bd;
    a.scale = j.scale;
    Step s = make_step(k, a, (count + k->tpb - 1) / k->tpb, 1, var == VAR_ROW ? "row" : "col");
    s.in_role = j.in_role;
    s.out_role = j.out_role;
    s.in_off = j.in_off;
    s.out_off = j.out_off;
    push_sliced(p.steps[di], s, slices, in_bd * es, out_bd * es);
    if (!algo.empty()) algo += "+";
    algo += "lds" + std::to_string(j.n) + (var == VAR_ROW ? "r" : "c");
    return AFFT_OK;
}

// Four-step for n = n1*n2 along a unit-stride axis (in_nd == out_nd == 1):
//   A: columns of the n1 x n2 view (stride n2), twiddle W_n^{k1*n2} fused,
//      into tmp at the same positions;
//   B: rows of tmp (length n2), written transposed: X[k1 + n1*k2].
// Intermediate bytes per four-step chunk; AFFT_FOURSTEP_CHUNK_MB overrides
// it (0: the whole batch at once).
static size_t fourstep_chunk_bytes() {
    static const size_t v = [] {
        const char* e = std::getenv("AFFT_FOURSTEP_CHUNK_MB");
        if (!e || !*e) return size_t(128) << 20;
        const long long mb = std::atoll(e);
        return mb > 0 ? size_t(mb) << 20 : ~size_t(0) >> 1;
    }();
    return v;
}

// The four-step split n1 (column length) of a power-of-two length 2^lg: the
// balanced one, except where the gfx950 tuning records show another one
// faster by more than 5 % on average over their shapes of that length:
// complex64 2^20 as 512 x 2048 (7 %), 2^23 as 4096 x 2048 (44 %), 2^24 as
// 8192 x 2048 (7 %); complex128 2^18 as 1024 x 256 (13 %), 2^19 as
// 256 x 2048 (10 %), 2^23 as 4096 x 2048 (32 %).
static int64_t fourstep_split(int prec, int lg) {
    static constexpr int tuned[2][25] = {
        // complex64, by lg
        {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 9, 0, 0, 12, 13},
        // complex128
        {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 10, 8, 0, 0, 0, 12, 0},
    };
    const int t = lg >= 0 && lg < 25 ? tuned[prec ? 1 : 0][lg] : 0;
    return 1LL << (t ? t : lg / 2);
}

static int emit_fourstep(Plan& p, int di, const AxisJob& j, std::string& algo) {
    const int prec = p.desc.precision;
    const int dir = di ? 1 : -1;
    const int es = prec ? 16 : 8;
    if (j.in_nd != 1 || j.out_nd != 1 || !is_pow2(j.n))
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "four-step needs a unit-stride power-of-two axis");
    const int lg = ilog2(j.n);
    const bool mf = p.desc.flags & AFFT_FLAG_MFMA;
    int64_t n1 = fourstep_split(prec, lg);
    if (mf) {
        // both MFMA passes have lengths in [256, 4096]: the balanced split
        // (n1 <= n2) unless a split is forced
        n1 = j.split ? j.split : (int64_t)1 << (lg / 2);
        if (!is_pow2(n1) || n1 < 256 || n1 > 4096 || j.n / n1 < 256 || j.n / n1 > 4096 || prec != 0)
            return set_error(AFFT_ERR_NOT_IMPLEMENTED,
                             "the MFMA form covers complex64 rows of 2^8 .. 2^24 points");
    } else if (j.split) {
        if (!is_pow2(j.split) || j.split < 2 || j.split >= j.n)
            return set_error(AFFT_ERR_NOT_IMPLEMENTED, "four-step split must be a power of two below n");
        n1 = j.split;
    }
    const int64_t n2 = j.n / n1;
    if (n2 > max_single_pass(prec) || n1 > max_single_pass(prec))
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "transform too large for this split");
    const KernelInfo* ka = find_kernel(prec, (int)n1, VAR_4A, dir);
    const KernelInfo* kb = find_kernel(prec, (int)n2, VAR_4B, dir);
    if (!ka || !kb) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "missing four-step kernel");
    // per-block transforms and launch shape of the two passes
    int ga = ka->tpb, gb = kb->tpb, ta = ka->threads, tb = kb->threads;
    const void* fa = ka->fn;
    const void* fb = kb->fn;
    if (mf) {
        fa = mfma_kernel((int)n1, VAR_4A, dir, &ga, &ta);
        fb = mfma_kernel((int)n2, VAR_4B, dir, &gb, &tb);
        if (!fa || !fb) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "missing MFMA four-step kernel");
    }
    // Transforms run in chunks of at most fourstep_chunk_bytes() of intermediate
    // data: pass A of a chunk writes TMP0 and pass B reads it back while much
    // of it is still in the 256 MB last-level cache.  Measured on 512 MiB
    // batches: 128 MiB chunks are 0-9 % faster than one pass over the whole
    // batch; 16-64 MiB chunks are slower (too few blocks per launch).
    const int64_t chunk = std::max<int64_t>(1, (int64_t)(fourstep_chunk_bytes() / (size_t)(j.n * es)));
    const int64_t cc = std::min<int64_t>(chunk, j.count);
    const int64_t need = cc * j.n * es;
    if ((size_t)need > p.tmp_bytes[0]) p.tmp_bytes[0] = (size_t)need;
    int rc;
    KArgs a{};
    if ((rc = schedule_twiddles(p, ka, &a.tw))) return rc;
    if ((rc = split_twiddles(p, prec, j.n, &a.tw4_shift, &a.tw4_lo, &a.tw4_hi))) return rc;
    a.tw4_L = j.n;
    a.count = n2;
    a.in_cd = 1;
    a.in_nd = n2;
    a.out_cd = 1;
    a.out_nd = n2;
    a.out_bd = j.n;
    a.scale = 1.0;
    KArgs b{};
    if ((rc = schedule_twiddles(p, kb, &b.tw))) return rc;
    b.count = n1;
    b.in_cd = n2;
    b.in_nd = 1;
    b.in_bd = j.n;
    b.out_cd = 1;
    b.out_nd = n1;
    b.scale = j.scale;
    // the transforms (c, y): c-major inside each y slice, cc at a time; a
    // launch covers its chunk's c with blockIdx.y
    for (int64_t y = 0; y < j.slices; ++y) {
        for (int64_t c0 = 0; c0 < j.count; c0 += cc) {
            const int64_t cnt = std::min(cc, j.count - c0);
            KArgs aa = a, bb = b;
            aa.in_bd = j.in_cd;
            bb.out_bd = j.out_cd;
            Step sa = make_step(ka, aa, (n2 + ga - 1) / ga, 1, "4stepA");
            sa.fn = fa;
            sa.block = dim3((unsigned)ta, 1, 1);
            sa.in_role = j.in_role;
            sa.in_off = j.in_off + (y * j.in_bd + c0 * j.in_cd) * es;
            sa.out_role = ROLE_TMP0;
            sa.out_off = 0;
            push_sliced(p.steps[di], sa, cnt, j.in_cd * es, j.n * es);
            Step sb = make_step(kb, bb, (n1 + gb - 1) / gb, 1, "4stepB");
            sb.fn = fb;
            sb.block = dim3((unsigned)tb, 1, 1);
            sb.in_role = ROLE_TMP0;
            sb.in_off = 0;
            sb.out_role = j.out_role;
            sb.out_off = j.out_off + (y * j.out_bd + c0 * j.out_cd) * es;
            push_sliced(p.steps[di], sb, cnt, j.n * es, j.out_cd * es);
        }
    }
    if (!algo.empty()) algo += "+";
    algo += (mf ? "mfma_fourstep" : "fourstep") + std::to_string(n1) + "x" + std::to_string(n2);
    return AFFT_OK;
}

// ---- runtime-size mixed radix ------------------------------------------------
const void* gen_kernel(int prec, int var, int dir);
const void* gen_ip_kernel(int prec, int dir);
const void* gen_ipg_kernel(int prec, int var, int dir);
const void* copy_kernel(int prec);

// Radix schedule for a 2/3/5/7-smooth n: radix-16 passes and the
// power-of-two remainder, with the 3s, 5s and 7s first when there are any.
// Empty if n has another prime factor.
std::vector<int> smooth_schedule(int64_t n) {
    std::vector<int> r;
    int64_t m = n;
    int a = 0;
    while (m % 2 == 0) { m /= 2; ++a; }
    std::vector<int> odd;
    for (int q : {3, 5, 7})
        while (m % q == 0) { m /= q; odd.push_back(q); }
    if (m != 1) return {};
    if (odd.empty()) {
        // powers of two: the remainder radix first (the compile-time
        // schedules' order, whose twiddle layout the ragged cores share)
        while (a >= 4) { r.push_back(16); a -= 4; }
        if (a > 0) r.insert(r.begin(), 1 << a);
        return r;
    }
    // Other smooth lengths: odd radices first, then the power-of-two
    // radices.  A radix-R pass writes its outputs R apart while Ns is small;
    // odd strides spread over the LDS banks, strides of 8 or 16 points
    // conflict, so the large power-of-two radices go where Ns is large.
    // Measured on 512 MiB batches: N=6000 0.431 -> 0.386 ms, N=100000
    // 1.044 -> 0.987 ms.
    for (int q : odd) r.push_back(q);
    if (a % 4) r.push_back(1 << (a % 4));
    for (int i = 0; i < a / 4; ++i) r.push_back(16);
    return r;
}

int gen_cap(int prec) { return prec ? 4096 : 8192; }
int gen_row_cap(int prec) { return prec ? 8192 : 16384; }

// Threads resident per CU for blocks of `threads` threads and `lds` bytes
// of LDS, at `wps` waves per SIMD (the kernel's register use): blocks
// counted by LDS (160 KiB) and by wave slots.
static int64_t gen_resident(int threads, int64_t lds, int wps) {
    const int64_t lds_blocks = std::max<int64_t>(1, (int64_t)(160 << 10) / std::max<int64_t>(1, lds));
    const int64_t wave_blocks = std::max(1, wps / ((threads / 64 + 3) / 4));
    return std::min(lds_blocks, wave_blocks) * threads;
}

// Threads of an in-place row block: the multiple of 64 from n/16 (each pass
// holds at most 16 points of a thread) up to the kernel's maximum that keeps
// the most threads resident per CU; on a tie the smaller block.  E.g.
// complex64 N=8000 (64 KiB): 512 threads, two blocks (0.334 ms) against 704
// threads, one block (0.481 ms).
static int gen_ip_threads(int64_t n, int prec, int64_t* resident = nullptr) {
    const int64_t es = prec ? 16 : 8;
    const int maxt = prec ? GenIP<double>::THREADS : GenIP<float>::THREADS;
    const int wps = prec ? GenIP<double>::WAVES_PER_SIMD : GenIP<float>::WAVES_PER_SIMD;
    int best = maxt;
    int64_t best_res = -1;
    for (int tt = (int)(((n + 15) / 16 + 63) / 64 * 64); tt <= maxt; tt += 64) {
        const int64_t res = gen_resident(tt, n * es, wps);
        if (res > best_res) { best_res = res; best = tt; }
    }
    if (resident) *resident = best_res;
    return best;
}

int64_t next_pow2(int64_t n) {
    int64_t m = 1;
    while (m < n) m <<= 1;
    return m;
}

struct GenCfg {
    int tpt, g, threads;
    int lds;
    bool ipg = false;   // k_gen_ipg: G transforms per block, passes in place in one buffer
};

// Two LDS buffers of G transforms, TPT threads each.  Row layouts: about 8
// points per thread up to 256 threads, and one transform per block above
// 4096 points: 1024 threads for complex64 (the LDS of two buffers allows one
// block per CU, so its waves are all the latency hiding there is), 512 for
// complex128.  Column layouts keep about 16 points per thread: more adjacent
// columns per block make longer coalesced segments.  Measured on 512 MiB
// batches: N=1000 rows 0.37 -> 0.26 ms, N=6000 rows 0.65 -> 0.50 ms (512
// threads) -> 0.45 ms (1024 threads); column passes lose with fewer columns
// per block.
static GenCfg gen_cfg_2b(int64_t n, int prec, bool rows) {
    GenCfg c;
    if (rows && n > 4096) c.tpt = prec == 0 ? GEN_ROWTHREADS : GEN_MAXTHREADS;
    else if (rows) c.tpt = (int)std::min<int64_t>(GEN_THREADS, next_pow2((n + 7) / 8));
    else c.tpt = (int)std::min<int64_t>(GEN_THREADS, next_pow2((n + 15) / 16));
    c.threads = std::max(GEN_THREADS, c.tpt);
    c.g = c.threads / c.tpt;
    c.lds = (int)(2 * c.g * n * (prec ? 16 : 8));
    return c;
}

// Contiguous rows run one transform per block with every pass in place in
// one LDS buffer (k_gen_ip) above the two-buffer kernel's reach (8192
// complex64 / 4096 complex128 points), and below it where that keeps at
// least twice the threads resident per CU of the two-buffer layout, whose
// rows hold 2n points (k_gen rows take 94-97 VGPRs for complex64, 4 waves
// per SIMD, and 165 for complex128, 3), or as many where the two-buffer
// layout runs a single block per CU.  Measured on 512 MiB batches
// (two-buffer -> in place): complex64 N=3500 0.416 -> 0.313 ms, N=4000
// 0.409 -> 0.323 ms (two blocks of 256 threads against four), N=4500 0.445
// -> 0.417 ms (one block of 1024 threads against two of 512), N=6000 0.387
// -> 0.365 ms, N=8000 0.349 -> 0.334 ms; complex128 N=3000 0.489 -> 0.328
// ms.  Kept on the two-buffer kernel: complex64 N=3000 (three blocks of 256
// against four: 0.341 against 0.350 ms), N=2000 (0.265 against 0.396 ms),
// complex128 N=2500 (0.290 against 0.335 ms).
static bool gen_use_ip(int64_t n, int prec) {
    if (n > gen_cap(prec)) return true;
    const GenCfg c = gen_cfg_2b(n, prec, true);
    int64_t res_ip = 0;
    gen_ip_threads(n, prec, &res_ip);
    const int64_t res_2b = gen_resident(c.threads, c.lds, prec ? 3 : 4);
    return res_ip >= 2 * res_2b || (res_ip >= res_2b && res_2b == c.threads);
}

// Whether one runtime-radix launch holds a transform of n points on this
// layout.
static bool gen_single_fits(int prec, int64_t n, bool rows) {
    return n <= gen_cap(prec) || (rows && n <= gen_row_cap(prec));
}

// With ipg, column layouts (both four-step passes, strided axes of N-D
// plans) run in place in one LDS buffer (k_gen_ipg) when they hold at most
// 16 points per thread: half the LDS of the two-buffer kernel, so twice the
// blocks per CU.  Measured on 512 MiB batches: complex64 N=100000 (250 x
// 400) 0.990 -> 0.831 ms, N=250000 (500 x 500) 1.012 -> 0.829 ms,
// complex128 N=100000 0.952 -> 0.700 ms.  Strided axes gain once blocks
// grow to 512 threads (below): complex64 1200 x 1500 1.325 -> 0.812 ms,
// 1000 x 1000 0.791 -> 0.570 ms, 250 x 250 x 64 1.179 -> 0.860 ms;
// complex128 1200 x 1500 0.752 -> 0.656 ms, 1000 x 1000 0.755 -> 0.618 ms
// (at 256 threads the in-place form had lost 2-7 % on them).
static GenCfg gen_cfg(int64_t n, int prec, bool rows, bool ipg = false) {
    if (rows && gen_use_ip(n, prec)) {
        GenCfg c;
        c.tpt = c.threads = gen_ip_threads(n, prec);
        c.g = 1;
        c.lds = (int)(n * (prec ? 16 : 8));
        return c;
    }
    GenCfg c = gen_cfg_2b(n, prec, rows);
    if (ipg && !rows && n <= 16 * (int64_t)c.tpt) {
        c.ipg = true;
        // Blocks of 512 threads where 256 would hold fewer than 16
        // (complex64) / 4 (complex128) transforms: longer column segments.
        // Measured on 512 MiB batches (256 -> 512 threads), complex64:
        // N=36000 (32 x 1125) 1.45 -> 0.85 ms, N=10^6 (1000 x 1000) 1.22 ->
        // 0.72 ms, N=100000 0.82 -> 0.71 ms, but N=24000 (64 x 375) 0.60 ->
        // 0.62 ms and N=162000 (400 x 405) 0.83 -> 0.88 ms; complex128
        // N=36000 0.84 -> 0.69 ms, while N=10^6 (4 transforms a block at 256)
        // went 0.71 -> 0.78 ms and N=100000 0.70 -> 0.77 ms.
        if (c.g < (prec ? 4 : 16))
            c.threads = std::min(2 * c.threads, prec ? GenIPG<double>::THREADS : GenIPG<float>::THREADS);
        c.g = c.threads / c.tpt;
        c.lds = (int)(c.g * n * (prec ? 16 : 8));
    }
    return c;
}

// ipg: may choose the in-place grouped kernel (four-step passes); *ipg says
// whether it did, and lds is then that kernel's.
static int gen_args(Plan& p, int64_t n, bool rows, GArgs& g, int* tpt, int* lds, int* threads,
                    bool* ipg = nullptr) {
    std::vector<int> rad = smooth_schedule(n);
    if (rad.empty() || (int)rad.size() > GEN_MAXPASS) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "no radix schedule");
    KernelInfo fake{};
    fake.prec = p.desc.precision;
    // reuse the schedule twiddle generator with the runtime radix list
    std::vector<double> tw;
    int64_t ns = rad[0];
    g.npass = (int)rad.size();
    g.ns[0] = 1;
    g.nsm[0] = mod_magic(1);
    g.twoff[0] = 0;
    g.rad[0] = rad[0];
    for (size_t i = 1; i < rad.size(); ++i) {
        const int R = rad[i];
        g.rad[i] = R;
        g.ns[i] = (int)ns;
        g.nsm[i] = mod_magic((unsigned)ns);
        g.twoff[i] = (int)(tw.size() / 2);
        // k-major (entry k*(R-1) + r-1): the runtime passes read a butterfly's
        // R-1 twiddles at compile-time offsets from one address
        for (int64_t kk = 0; kk < ns; ++kk)
            for (int r = 1; r < R; ++r) {
                double re, im;
                root_of_unity(kk * r, ns * R, &re, &im);
                tw.push_back(re);
                tw.push_back(im);
            }
        ns *= R;
    }
    int rc;
    if (p.desc.precision == 0) {
        std::vector<float> f(tw.begin(), tw.end());
        rc = upload_table(p, f.data(), f.size() * 4, &g.k.tw);
    } else {
        rc = upload_table(p, tw.data(), tw.size() * 8, &g.k.tw);
    }
    if (rc) return rc;
    GenCfg c = gen_cfg(n, p.desc.precision, rows, ipg != nullptr);
    g.n = (int)n;
    g.tpt = c.tpt;
    *tpt = c.tpt;
    *lds = c.lds;
    *threads = c.threads;
    if (ipg) *ipg = c.ipg;
    return AFFT_OK;
}

static Step make_gen_step(const void* fn, const GArgs& g, int64_t gx, int lds, int threads, const std::string& name) {
    Step s;
    s.fn = fn;
    s.grid = dim3((unsigned)gx, 1, 1);
    s.block = dim3((unsigned)threads, 1, 1);
    s.lds_bytes = lds;
    s.arg.resize(sizeof(GArgs));
    std::memcpy(s.arg.data(), &g, sizeof(GArgs));
    s.in_field = (int)(offsetof(GArgs, k) + offsetof(KArgs, in));
    s.out_field = (int)(offsetof(GArgs, k) + offsetof(KArgs, out));
    s.nt_field = (int)(offsetof(GArgs, k) + offsetof(KArgs, nt));
    s.name = name;
    return s;
}

static int prep_dynamic_lds(const void* fn, int lds) {
    if (lds > 65536) AFFT_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
    return AFFT_OK;
}

static int emit_gen_single(Plan& p, int di, const AxisJob& j, std::string& algo) {
    const int prec = p.desc.precision;
    const int dir = di ? 1 : -1;
    const int es = prec ? 16 : 8;
    if (p.desc.flags & AFFT_FLAG_MFMA) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "mixed radix has no MFMA form");
    int var;
    int64_t count = j.count, slices = j.slices, in_bd = j.in_bd, out_bd = j.out_bd;
    if (j.in_nd == 1 && j.out_nd == 1) var = VAR_ROW;
    else if (j.in_cd == 1 && j.out_cd == 1) var = VAR_COL;
    else {
        var = VAR_COL;
        if (count != 1) {
            if (slices != 1) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "layout with no unit stride and two batch axes");
            slices = count;
            in_bd = j.in_cd;
            out_bd = j.out_cd;
            count = 1;
        }
    }
    if (!gen_single_fits(prec, j.n, var == VAR_ROW))
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "mixed radix: length too large for one pass on this layout");
    GArgs g{};
    int tpt, lds, threads, rc;
    bool ipg = false;
    // strided columns above 128 points (16 threads or more a column) on the
    // in-place kernel; below, its separate store phase costs more than the
    // LDS it saves (the 3-point axis of 3 x 5000 x 7: 2.7 % slower)
    if ((rc = gen_args(p, j.n, var == VAR_ROW, g, &tpt, &lds, &threads, var == VAR_COL && j.n > 128 ? &ipg : nullptr)))
        return rc;
    const bool ip = var == VAR_ROW && gen_use_ip(j.n, prec);
    const void* fn = ip ? gen_ip_kernel(prec, dir) : ipg ? gen_ipg_kernel(prec, var, dir) : gen_kernel(prec, var, dir);
    if ((rc = prep_dynamic_lds(fn, lds))) return rc;
    g.k.count = count;
    g.k.in_cd = j.in_cd;
    g.k.out_cd = j.out_cd;
    g.k.in_nd = j.in_nd;
    g.k.out_nd = j.out_nd;
    g.k.in_bd = in_bd;
    g.k.out_bd = out_bd;
    g.k.scale = j.scale;
    const int G = threads / tpt;
    Step s = make_gen_step(fn, g, (count + G - 1) / G, lds, threads, "gen" + std::to_string(j.n));
    s.in_role = j.in_role;
    s.out_role = j.out_role;
    s.in_off = j.in_off;
    s.out_off = j.out_off;
    push_sliced(p.steps[di], s, slices, in_bd * es, out_bd * es);
    if (!algo.empty()) algo += "+";
    algo += "mixed" + std::to_string(j.n) + (var == VAR_ROW ? "r" : "c");
    return AFFT_OK;
}

// Default split n = n1 x n2 of a smooth four-step: the divisor pair of least
// estimated cost, each factor a power of two held in LDS (the compile-time
// kernel) or a runtime-radix length of 8 to 4096 points (the in-place
// column kernel, at most 16 points per thread), with n1/sqrt(n) in
// [0.2, 5].  A factor costs 3 on the compile-time kernel, else one per pass
// plus one per radix-2 and a half per radix-4 pass; ties go to the more
// balanced pair, then the smaller n1.  Fitted on 33 complex64 / complex128
// lengths from 9000 to 10^6 with every such split timed on 512 MiB batches:
// 5.4 % mean regret, 29 % at worst (5.8 % and 6.0 % on either half fitted
// on the other), against 15.6 % and 71 % for the earlier rule (the
// power-of-two part as the column
// factor when 16..4096 points with the rest at most 2048, else the most
// balanced divisors): complex64 18000 as 16 x 1125 took 0.99 ms against
// 0.59 ms as 225 x 80, 30000 as 16 x 1875 1.06 against 0.62 ms.  Lengths
// with no such pair take the most balanced divisors within the caps.
int64_t gen_fourstep_split(int64_t n, int prec) {
    const int64_t cap = gen_cap(prec), rcap = std::min<int64_t>(cap, 4096);
    auto usable = [&](int64_t m) { return is_pow2(m) ? m <= max_single_pass(prec) : m <= rcap; };
    auto cost = [&](int64_t m) {
        if (is_pow2(m)) return 3.0;
        const std::vector<int> r = smooth_schedule(m);
        double c = (double)r.size();
        for (int q : r) c += q == 2 ? 1.0 : q == 4 ? 0.5 : 0.0;
        return c;
    };
    const double rt = std::sqrt((double)n);
    int64_t n1 = 0;
    double best = 0, bal = 0;
    const int64_t lo = std::max<int64_t>(8, (int64_t)std::ceil(0.2 * rt)), hi = std::min<int64_t>(n / 8, (int64_t)(5.0 * rt));
    for (int64_t d = lo; d <= hi; ++d) {
        if (n % d || !usable(d) || !usable(n / d)) continue;
        const double c = cost(d) + cost(n / d), b = std::fabs(std::log(d / rt));
        if (!n1 || c < best - 1e-9 || (c < best + 1e-9 && b < bal - 1e-9)) {
            n1 = d;
            best = c;
            bal = b;
        }
    }
    if (!n1)
        for (int64_t d = 1; d * d <= n; ++d)
            if (n % d == 0 && d <= cap && n / d <= cap) n1 = d;
    return n1;
}

// Four-step over runtime kernels for smooth n beyond one LDS pass.
static int emit_gen_fourstep(Plan& p, int di, const AxisJob& j, std::string& algo) {
    const int prec = p.desc.precision;
    const int dir = di ? 1 : -1;
    const int es = prec ? 16 : 8;
    if (j.in_nd != 1 || j.out_nd != 1) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "four-step needs unit stride");
    const int cap = gen_cap(prec);
    int64_t n1 = gen_fourstep_split(j.n, prec);
    if (j.split) {
        // a power-of-two factor may use the compile-time kernels' larger cap
        auto fits = [&](int64_t f) { return f <= (is_pow2(f) ? (int64_t)max_single_pass(prec) : (int64_t)cap); };
        if (j.split < 2 || j.n % j.split || !fits(j.split) || !fits(j.n / j.split))
            return set_error(AFFT_ERR_NOT_IMPLEMENTED, "four-step split must divide n into LDS-sized factors");
        n1 = j.split;
    }
    if (!n1) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "no four-step split for n=" + std::to_string(j.n));
    const int64_t n2 = j.n / n1;
    // chunks of transforms whose intermediate data stays in the last-level cache
    const int64_t cc = std::min<int64_t>(j.count, std::max<int64_t>(1, (int64_t)(fourstep_chunk_bytes() / (size_t)(j.n * es))));
    const size_t need = (size_t)(cc * j.n * es);
    if (need > p.tmp_bytes[0]) p.tmp_bytes[0] = need;
    int rc;
    // A power-of-two factor held in LDS runs on the compile-time Stockham
    // kernel of that length, the other factor on the runtime passes.
    const KernelInfo* ka = (is_pow2(n1) && n1 >= 2) ? find_kernel(prec, (int)n1, VAR_4A, dir) : nullptr;
    const KernelInfo* kb = (is_pow2(n2) && n2 >= 2) ? find_kernel(prec, (int)n2, VAR_4B, dir) : nullptr;
    GArgs ga{}, gb{};
    int tpa = 1, ldsa = 0, tha = 1, tpb = 1, ldsb = 0, thb = 1;
    bool ipa = false, ipb = false;
    if (ka) {
        if ((rc = schedule_twiddles(p, ka, &ga.k.tw))) return rc;
    } else if ((rc = gen_args(p, n1, false, ga, &tpa, &ldsa, &tha, &ipa))) {
        return rc;
    }
    if (kb) {
        if ((rc = schedule_twiddles(p, kb, &gb.k.tw))) return rc;
    } else if ((rc = gen_args(p, n2, false, gb, &tpb, &ldsb, &thb, &ipb))) {
        return rc;
    }
    if ((rc = split_twiddles(p, prec, j.n, &ga.k.tw4_shift, &ga.k.tw4_lo, &ga.k.tw4_hi))) return rc;
    const void* fa = ka ? ka->fn : ipa ? gen_ipg_kernel(prec, VAR_4A, dir) : gen_kernel(prec, VAR_4A, dir);
    const void* fb = kb ? kb->fn : ipb ? gen_ipg_kernel(prec, VAR_4B, dir) : gen_kernel(prec, VAR_4B, dir);
    if ((!ka && (rc = prep_dynamic_lds(fa, ldsa))) || (!kb && (rc = prep_dynamic_lds(fb, ldsb)))) return rc;
    ga.k.tw4_L = j.n;
    ga.k.count = n2;
    ga.k.in_cd = 1; ga.k.in_nd = n2; ga.k.out_cd = 1; ga.k.out_nd = n2;
    ga.k.out_bd = j.n;
    ga.k.scale = 1.0;
    gb.k.count = n1;
    gb.k.in_cd = n2; gb.k.in_nd = 1; gb.k.in_bd = j.n;
    gb.k.out_cd = 1; gb.k.out_nd = n1;
    gb.k.scale = j.scale;
    const int Ga = tha / tpa, Gb = thb / tpb;
    for (int64_t y = 0; y < j.slices; ++y) {
        for (int64_t c0 = 0; c0 < j.count; c0 += cc) {
            const int64_t cnt = std::min(cc, j.count - c0);
            GArgs a2 = ga, b2 = gb;
            a2.k.in_bd = j.in_cd;
            b2.k.out_bd = j.out_cd;
            Step sa = ka ? make_step(ka, a2.k, (n2 + ka->tpb - 1) / ka->tpb, 1, "4stepA")
                         : make_gen_step(fa, a2, (n2 + Ga - 1) / Ga, ldsa, tha, "gen4A");
            sa.in_role = j.in_role;
            sa.in_off = j.in_off + (y * j.in_bd + c0 * j.in_cd) * es;
            sa.out_role = ROLE_TMP0;
            sa.out_off = 0;
            push_sliced(p.steps[di], sa, cnt, j.in_cd * es, j.n * es);
            Step sb = kb ? make_step(kb, b2.k, (n1 + kb->tpb - 1) / kb->tpb, 1, "4stepB")
                         : make_gen_step(fb, b2, (n1 + Gb - 1) / Gb, ldsb, thb, "gen4B");
            sb.in_role = ROLE_TMP0;
            sb.in_off = 0;
            sb.out_role = j.out_role;
            sb.out_off = j.out_off + (y * j.out_bd + c0 * j.out_cd) * es;
            push_sliced(p.steps[di], sb, cnt, j.n * es, j.out_cd * es);
        }
    }
    if (!algo.empty()) algo += "+";
    algo += "mixed4step" + std::to_string(n1) + "x" + std::to_string(n2);
    return AFFT_OK;
}

// ---- Bluestein ---------------------------------------------------------------
// In-place iterative radix-2 FFT in long double: builds the Bluestein filter
// spectrum at plan time.
static void host_fft_ld(std::vector<long double>& re, std::vector<long double>& im) {
    const size_t n = re.size();
    for (size_t i = 1, jj = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; jj & bit; bit >>= 1) jj ^= bit;
        jj ^= bit;
        if (i < jj) { std::swap(re[i], re[jj]); std::swap(im[i], im[jj]); }
    }
    // twiddles W_n^k, k < n/2, computed once
    std::vector<long double> wr(n / 2 + 1), wi(n / 2 + 1);
    for (size_t k = 0; k < n / 2; ++k) {
        double c, s;
        root_of_unity((int64_t)k, (int64_t)n, &c, &s);
        wr[k] = c;
        wi[k] = s;
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const size_t half = len / 2, step = n / len;
        for (size_t b = 0; b < n; b += len) {
            for (size_t k = 0; k < half; ++k) {
                const long double cr = wr[k * step], ci = wi[k * step];
                const size_t u = b + k, v = b + k + half;
                const long double xr = re[v] * cr - im[v] * ci, xi = re[v] * ci + im[v] * cr;
                re[v] = re[u] - xr; im[v] = im[u] - xi;
                re[u] += xr; im[u] += xi;
            }
        }
    }
}

// Forward spectrum of one sequence of m = 2^k complex128 points (interleaved
// re, im in `v`, replaced by its DFT), on the plan's device with a complex128
// plan of that length forced onto the power-of-two routes (never Bluestein).
static int device_fft_c128(Plan& p, std::vector<double>& v, int64_t m) {
    afft_plan_desc d{};
    d.kind = AFFT_C2C;
    d.precision = AFFT_COMPLEX128;
    d.rank = 1;
    d.dims[0] = m;
    d.batch = 1;
    d.strategy = m <= max_single_pass(1) ? AFFT_STRATEGY_STOCKHAM : AFFT_STRATEGY_SIXSTEP;
    afft_plan sub = nullptr;
    int rc = afft_plan_create_desc(static_cast<afft_context>(p.ctx), &d, &sub);
    if (rc) return rc;
    const size_t bytes = (size_t)m * 16;
    void* buf = nullptr;
    hipError_t e = hipMalloc(&buf, 2 * bytes);
    if (e == hipSuccess) e = hipMemcpy(buf, v.data(), bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        rc = exec_plan(*sub, AFFT_FORWARD, (char*)buf + bytes, buf, sub->stream);
        if (rc == AFFT_OK) e = hipStreamSynchronize(sub->stream);
        if (rc == AFFT_OK && e == hipSuccess) e = hipMemcpy(v.data(), (char*)buf + bytes, bytes, hipMemcpyDeviceToHost);
    }
    if (buf) (void)hipFree(buf);
    afft_plan_destroy(sub);
    if (rc) return rc;
    if (e != hipSuccess) return set_error(AFFT_ERR_DEVICE, std::string("Bluestein filter spectrum: ") + hipGetErrorString(e));
    return AFFT_OK;
}

// The chirp exp(-iπ k²/n) = W_{2n}^{k² mod 2n}, k < n, in FP64.  Above 4096
// points from two long-double tables of about sqrt(2n) roots each,
// W_{2n}^r = W_{2n}^{(r >> s) 2^s} W_{2n}^{r & (2^s - 1)}, multiplied in long
// double and rounded once (the product's error is far below a double ulp).
// AFFT_FLAG_REF_CHIRP uses the reference's unreduced argument π k²/n.
static void bluestein_chirp(int64_t n, bool refchirp, std::vector<double>& ch) {
    ch.assign(2 * n, 0.0);
    if (refchirp) {
        for (int64_t k = 0; k < n; ++k) {
            const double a = -3.14159265358979323846 * (double)(k * k) * (1.0 / (double)n);
            ch[2 * k] = std::cos(a);
            ch[2 * k + 1] = std::sin(a);
        }
        return;
    }
    const int64_t L = 2 * n;
    if (n <= 4096) {
        for (int64_t k = 0, r = 0; k < n; r = (r + 2 * k + 1) % L, ++k)
            root_of_unity(r, L, &ch[2 * k], &ch[2 * k + 1]);   // r = k^2 mod 2n
        return;
    }
    int sh = 0;
    while ((int64_t(1) << (2 * sh)) < L) ++sh;
    const int64_t B = int64_t(1) << sh, H = (L + B - 1) / B;
    std::vector<long double> lr(B), li(B), hr(H), hi(H);
    for (int64_t i = 0; i < B; ++i) root_of_unity_ld(i, L, &lr[i], &li[i]);
    for (int64_t i = 0; i < H; ++i) root_of_unity_ld(i * B, L, &hr[i], &hi[i]);
    for (int64_t k = 0, r = 0; k < n; r = (r + 2 * k + 1) % L, ++k) {
        const int64_t a = r & (B - 1), b = r >> sh;
        ch[2 * k] = (double)(lr[a] * hr[b] - li[a] * hi[b]);
        ch[2 * k + 1] = (double)(lr[a] * hi[b] + li[a] * hr[b]);
    }
}

int blue_tables(Plan& p, int64_t n, int64_t m, double inv_scale, BlueTables& bt, int64_t tr_m1) {
    const bool refchirp = p.desc.flags & AFFT_FLAG_REF_CHIRP;
    int rc;
    auto up = [&](const std::vector<double>& v, const void** dev) -> int {
        if (p.desc.precision == 0) {
            std::vector<float> f(v.begin(), v.end());
            return upload_table(p, f.data(), f.size() * 4, dev);
        }
        return upload_table(p, v.data(), v.size() * 8, dev);
    };
    // chirp and filter spectra are shared by the plan's two directions
    const auto key = std::make_tuple(n, m, tr_m1);
    auto it = p.blue_cache.find(key);
    if (it == p.blue_cache.end()) {
        BlueCache bc;
        bluestein_chirp(n, refchirp, bc.ch);
        // Forward filter b[k] = conj(chirp[k]), mirrored (b[m-k] = b[k]), zero
        // between.  Its spectrum B: a long-double host FFT up to 4096 points,
        // above that one complex128 FFT on the device (each costs about
        // 0.1-0.5 ms around 4096 points).  The inverse filter is
        // conj(b), whose spectrum is conj(B[-k mod m]).
        std::vector<double> spec(2 * m, 0.0);
        spec[0] = bc.ch[0];
        spec[1] = -bc.ch[1];
        for (int64_t k = 1; k < n; ++k) {
            spec[2 * k] = spec[2 * (m - k)] = bc.ch[2 * k];
            spec[2 * k + 1] = spec[2 * (m - k) + 1] = -bc.ch[2 * k + 1];
        }
        if (m <= 4096) {
            std::vector<long double> re(m), im(m);
            for (int64_t k = 0; k < m; ++k) {
                re[k] = spec[2 * k];
                im[k] = spec[2 * k + 1];
            }
            host_fft_ld(re, im);
            for (int64_t k = 0; k < m; ++k) {
                spec[2 * k] = (double)re[k];
                spec[2 * k + 1] = (double)im[k];
            }
        } else if ((rc = device_fft_c128(p, spec, m))) {
            return rc;
        }
        std::vector<double> filt[2];
        // natural order, or row-major [k1][k2] of k = k1 + tr_m1 * k2
        const int64_t m2 = tr_m1 > 0 ? m / tr_m1 : 0;
        const double inv_m = 1.0 / (double)m;   // m is a power of two: exact
        for (int d = 0; d < 2; ++d) filt[d].resize(2 * m);
        for (int64_t k = 0; k < m; ++k) {
            const int64_t at = tr_m1 > 0 ? (k % tr_m1) * m2 + k / tr_m1 : k;
            const int64_t mk_ = (m - k) & (m - 1);
            filt[0][2 * at] = spec[2 * k] * inv_m;
            filt[0][2 * at + 1] = spec[2 * k + 1] * inv_m;
            filt[1][2 * at] = spec[2 * mk_] * inv_m;
            filt[1][2 * at + 1] = -spec[2 * mk_ + 1] * inv_m;
        }
        if ((rc = up(bc.ch, &bc.chirp))) return rc;
        for (int d = 0; d < 2; ++d)
            if ((rc = up(filt[d], &bc.filt[d]))) return rc;
        it = p.blue_cache.emplace(key, std::move(bc)).first;
    }
    const std::vector<double>& ch = it->second.ch;
    bt.chirp = it->second.chirp;
    bt.filt[0] = it->second.filt[0];
    bt.filt[1] = it->second.filt[1];
    // post-multiply tables: the chirp itself (forward), conj(chirp) times the
    // inverse scale (inverse)
    bt.post[0] = bt.chirp;
    std::vector<double> post(2 * n);
    for (int64_t k = 0; k < n; ++k) {
        post[2 * k] = ch[2 * k] * inv_scale;
        post[2 * k + 1] = -ch[2 * k + 1] * inv_scale;
    }
    return up(post, &bt.post[1]);
}

static int emit_bluestein(Plan& p, int di, const AxisJob& j, std::string& algo) {
    const int prec = p.desc.precision;
    const int dir = di ? 1 : -1;
    const int es = prec ? 16 : 8;
    if (j.in_nd != 1 || j.out_nd != 1) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "Bluestein needs unit stride");
    const int64_t n = j.n, m = next_pow2(2 * n - 1);
    BlueTables bt;
    // the inverse scale rides on the post-multiply table in the chained form
    int rc = blue_tables(p, n, m, di ? j.scale : 1.0, bt);
    if (rc) return rc;
    if (m <= max_single_pass(prec)) {
        const KernelInfo* k = find_kernel(prec, (int)m, VAR_BLUE, dir);
        if (!k) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "no Bluestein kernel for M=" + std::to_string(m));
        KArgs dummy{};
        int rc2 = schedule_twiddles(p, k, &dummy.tw);
        if (rc2) return rc2;
        BArgs b{};
        b.tw = dummy.tw;
        b.chirp = bt.chirp;
        b.filt = bt.filt[di];
        b.count = j.count;
        b.in_cd = j.in_cd;
        b.out_cd = j.out_cd;
        b.nlog = n;
        b.scale = j.scale;
        Step s;
        s.fn = k->fn;
        s.grid = dim3((unsigned)((j.count + k->tpb - 1) / k->tpb), 1, 1);
        s.block = dim3((unsigned)k->threads, 1, 1);
        s.arg.resize(sizeof(BArgs));
        std::memcpy(s.arg.data(), &b, sizeof(BArgs));
        s.in_field = (int)offsetof(BArgs, in);
        s.out_field = (int)offsetof(BArgs, out);
        s.in_role = j.in_role;
        s.out_role = j.out_role;
        s.in_off = j.in_off;
        s.out_off = j.out_off;
        s.name = "bluestein" + std::to_string(n);
        for (int64_t y = 0; y < j.slices; ++y) {
            Step t = s;
            t.in_off += y * j.in_bd * es;
            t.out_off += y * j.out_bd * es;
            p.steps[di].push_back(t);
        }
        if (!algo.empty()) algo += "+";
        algo += "bluestein" + std::to_string(n) + "m" + std::to_string(m);
        return AFFT_OK;
    }
    // chained form on the four-step kernels of M = m1*m2
    const int lg = ilog2(m);
    const int64_t m1 = 1LL << (lg / 2), m2 = m / m1;
    if (m2 > max_single_pass(prec)) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "Bluestein length too large");
    const KernelInfo* ka[2] = {find_kernel(prec, (int)m1, VAR_4A, -1), find_kernel(prec, (int)m1, VAR_4A, 1)};
    const KernelInfo* kb[2] = {find_kernel(prec, (int)m2, VAR_4B, -1), find_kernel(prec, (int)m2, VAR_4B, 1)};
    if (!ka[0] || !ka[1] || !kb[0] || !kb[1]) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "missing four-step kernel");
    // chunks of transforms whose two M-point intermediates stay in the
    // last-level cache between the four passes
    const int64_t cc = std::min<int64_t>(j.count, std::max<int64_t>(1, (int64_t)(fourstep_chunk_bytes() / (size_t)(2 * m * es))));
    const size_t need = (size_t)(cc * m * es);
    for (int t = 0; t < 2; ++t)
        if (need > p.tmp_bytes[t]) p.tmp_bytes[t] = need;
    KArgs A[2]{}, B[2]{};
    for (int d = 0; d < 2; ++d) {
        if ((rc = schedule_twiddles(p, ka[d], &A[d].tw))) return rc;
        if ((rc = schedule_twiddles(p, kb[d], &B[d].tw))) return rc;
        if ((rc = split_twiddles(p, prec, m, &A[d].tw4_shift, &A[d].tw4_lo, &A[d].tw4_hi))) return rc;
        A[d].tw4_L = m;
        A[d].count = m2;
        A[d].in_cd = 1; A[d].in_nd = m2; A[d].out_cd = 1; A[d].out_nd = m2;
        A[d].out_bd = m;
        A[d].scale = 1.0;
        B[d].count = m1;
        B[d].in_cd = m2; B[d].in_nd = 1; B[d].in_bd = m;
        B[d].out_cd = 1; B[d].out_nd = m1;
        B[d].scale = 1.0;
    }
    // A1: chirp-multiply + zero-pad fused into the first column pass
    A[0].fuse = FUSE_CHIRP_IN;
    A[0].fvec = bt.chirp;
    A[0].fconj = di;
    A[0].flen = n;
    // B1: pointwise filter product fused into the transposing store
    B[0].fuse = FUSE_MUL_OUT;
    B[0].fvec = bt.filt[di];
    B[0].flen = m;
    B[0].out_bd = m;
    // B2: post-chirp (and 1/N for the inverse), cropped to n points
    B[1].fuse = FUSE_CROP_OUT;
    B[1].fvec = bt.post[di];
    B[1].flen = n;
    for (int64_t y = 0; y < j.slices; ++y) {
        for (int64_t c0 = 0; c0 < j.count; c0 += cc) {
            const int64_t cnt = std::min(cc, j.count - c0);
            KArgs a1 = A[0];
            a1.in_bd = j.in_cd;
            Step s1 = make_step(ka[0], a1, (m2 + ka[0]->tpb - 1) / ka[0]->tpb, 1, "blueA1");
            s1.in_role = j.in_role; s1.in_off = j.in_off + (y * j.in_bd + c0 * j.in_cd) * es;
            s1.out_role = ROLE_TMP0; s1.out_off = 0;
            push_sliced(p.steps[di], s1, cnt, j.in_cd * es, m * es);
            Step s2 = make_step(kb[0], B[0], (m1 + kb[0]->tpb - 1) / kb[0]->tpb, 1, "blueB1");
            s2.in_role = ROLE_TMP0; s2.in_off = 0;
            s2.out_role = ROLE_TMP1; s2.out_off = 0;
            push_sliced(p.steps[di], s2, cnt, m * es, m * es);
            KArgs a2 = A[1];
            a2.in_bd = m;
            Step s3 = make_step(ka[1], a2, (m2 + ka[1]->tpb - 1) / ka[1]->tpb, 1, "blueA2");
            s3.in_role = ROLE_TMP1; s3.in_off = 0;
            s3.out_role = ROLE_TMP0; s3.out_off = 0;
            push_sliced(p.steps[di], s3, cnt, m * es, m * es);
            KArgs b2 = B[1];
            b2.out_bd = j.out_cd;
            Step s4 = make_step(kb[1], b2, (m1 + kb[1]->tpb - 1) / kb[1]->tpb, 1, "blueB2");
            s4.in_role = ROLE_TMP0; s4.in_off = 0;
            s4.out_role = j.out_role; s4.out_off = j.out_off + (y * j.out_bd + c0 * j.out_cd) * es;
            push_sliced(p.steps[di], s4, cnt, m * es, j.out_cd * es);
        }
    }
    if (!algo.empty()) algo += "+";
    algo += "bluestein" + std::to_string(n) + "m" + std::to_string(m1) + "x" + std::to_string(m2);
    return AFFT_OK;
}

// ---- copies (length-1 axes, gathers) ------------------------------------------
static Step make_copy(int prec, int64_t lines, int64_t len, int64_t icd, int64_t ind, int64_t ocd, int64_t ond,
                      double scale) {
    CopyArgs c{};
    c.lines = lines;
    c.len = len;
    c.in_cd = icd; c.in_nd = ind; c.out_cd = ocd; c.out_nd = ond;
    c.scale = scale;
    Step s;
    s.fn = copy_kernel(prec);
    const int64_t tot = lines * len;
    s.grid = dim3((unsigned)((tot + 255) / 256), 1, 1);
    s.block = dim3(256, 1, 1);
    s.arg.resize(sizeof(CopyArgs));
    std::memcpy(s.arg.data(), &c, sizeof(CopyArgs));
    s.in_field = (int)offsetof(CopyArgs, in);
    s.out_field = (int)offsetof(CopyArgs, out);
    s.name = "copy";
    return s;
}

static int emit_axis_direct(Plan& p, int di, const AxisJob& j, std::string& algo);

// Paths that need unit-stride rows get strided lines through a gather into
// TMP1, the row transform in place there, and a scatter back.
static int emit_via_gather(Plan& p, int di, const AxisJob& j, std::string& algo) {
    const int prec = p.desc.precision;
    const int es = prec ? 16 : 8;
    const int64_t lines = j.count * j.slices;
    const size_t need = (size_t)(lines * j.n * es);
    if (need > p.tmp_bytes[1]) p.tmp_bytes[1] = need;
    for (int64_t y = 0; y < j.slices; ++y) {
        Step g = make_copy(prec, j.count, j.n, j.in_cd, j.in_nd, j.n, 1, 1.0);
        g.in_role = j.in_role; g.in_off = j.in_off + y * j.in_bd * es;
        g.out_role = ROLE_TMP1; g.out_off = y * j.count * j.n * es;
        p.steps[di].push_back(g);
    }
    AxisJob r = j;
    r.count = lines;
    r.slices = 1;
    r.in_cd = r.out_cd = j.n;
    r.in_nd = r.out_nd = 1;
    r.in_bd = r.out_bd = 0;
    r.in_role = r.out_role = ROLE_TMP1;
    r.in_off = r.out_off = 0;
    int rc = emit_axis_direct(p, di, r, algo);
    if (rc) return rc;
    for (int64_t y = 0; y < j.slices; ++y) {
        Step s = make_copy(prec, j.count, j.n, j.n, 1, j.out_cd, j.out_nd, 1.0);
        s.in_role = ROLE_TMP1; s.in_off = y * j.count * j.n * es;
        s.out_role = j.out_role; s.out_off = j.out_off + y * j.out_bd * es;
        p.steps[di].push_back(s);
    }
    algo += "(gathered)";
    return AFFT_OK;
}

static bool needs_rows(const Plan& p, int64_t n, int strategy) {
    const int prec = p.desc.precision;
    if (strategy == AFFT_STRATEGY_SIXSTEP || strategy == AFFT_STRATEGY_BLUESTEIN) return true;
    if ((n & (n - 1)) == 0) return n > max_single_pass(prec);
    if (!smooth_schedule(n).empty()) return n > gen_cap(prec);
    return true;  // Bluestein
}

static int emit_axis_direct(Plan& p, int di, const AxisJob& j, std::string& algo) {
    const int prec = p.desc.precision;
    const bool smooth = !smooth_schedule(j.n).empty();
    const bool rows = j.in_nd == 1 && j.out_nd == 1;
    if ((p.desc.flags & AFFT_FLAG_MFMA) &&
        (!is_pow2(j.n) || j.strategy == AFFT_STRATEGY_MIXED || j.strategy == AFFT_STRATEGY_BLUESTEIN))
        return set_error(AFFT_ERR_NOT_IMPLEMENTED,
                         "the MFMA form covers powers of two on the Stockham and four-step routes only");
    switch (j.strategy) {
    case AFFT_STRATEGY_AUTO: break;
    case AFFT_STRATEGY_STOCKHAM:
        if (is_pow2(j.n) && j.n <= max_single_pass(prec)) return emit_single(p, di, j, algo);
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "Stockham strategy needs a power of two held in LDS");
    case AFFT_STRATEGY_MIXED:
        if (smooth && gen_single_fits(prec, j.n, rows)) return emit_gen_single(p, di, j, algo);
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "mixed-radix strategy needs a 7-smooth size held in LDS");
    case AFFT_STRATEGY_SIXSTEP:
        if (is_pow2(j.n)) return emit_fourstep(p, di, j, algo);
        if (smooth) return emit_gen_fourstep(p, di, j, algo);
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "four-step strategy needs a 7-smooth size");
    case AFFT_STRATEGY_BLUESTEIN:
        if (j.n >= 2) return emit_bluestein(p, di, j, algo);
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "Bluestein needs n >= 2");
    default:
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "unknown strategy");
    }
    if (is_pow2(j.n) && j.n <= max_single_pass(prec)) return emit_single(p, di, j, algo);
    if (is_pow2(j.n)) return emit_fourstep(p, di, j, algo);
    if (!smooth_schedule(j.n).empty()) {
        if (gen_single_fits(prec, j.n, rows)) return emit_gen_single(p, di, j, algo);
        return emit_gen_fourstep(p, di, j, algo);
    }
    return emit_bluestein(p, di, j, algo);
}

static int emit_axis(Plan& p, int di, const AxisJob& j, std::string& algo) {
    const int prec = p.desc.precision;
    const int es = prec ? 16 : 8;
    if (j.n == 1) {
        for (int64_t y = 0; y < j.slices; ++y) {
            Step s = make_copy(prec, j.count, 1, j.in_cd, 1, j.out_cd, 1, j.scale);
            s.in_role = j.in_role; s.in_off = j.in_off + y * j.in_bd * es;
            s.out_role = j.out_role; s.out_off = j.out_off + y * j.out_bd * es;
            p.steps[di].push_back(s);
        }
        if (!algo.empty()) algo += "+";
        algo += "copy";
        return AFFT_OK;
    }
    if (needs_rows(p, j.n, j.strategy) && (j.in_nd != 1 || j.out_nd != 1)) return emit_via_gather(p, di, j, algo);
    return emit_axis_direct(p, di, j, algo);
}

// Complex transforms over the axes `axes` (processed in the given order) of
// a row-major array with dims cd[0..rank-1]; batch copies at distances
// in_dist / out_dist (elements).  The first axis reads `in_role`, every axis
// writes `out_role`, later axes work in place there.
static int emit_nd_axes(Plan& p, int di, int rank, const int64_t* cd, const std::vector<int>& axes, int64_t batch,
                        int in_role, int64_t in_dist, int out_role, int64_t out_dist, std::string& algo) {
    int64_t total = 1;
    for (int i = 0; i < rank; ++i) total *= cd[i];
    bool first = true;
    for (int ax : axes) {
        const int64_t len = cd[ax];
        int64_t inner = 1;
        for (int i = ax + 1; i < rank; ++i) inner *= cd[i];
        const int64_t outer = total / (len * inner);
        if (len == 1 && !(first && in_role != out_role)) continue;   // identity axis
        AxisJob j{};
        j.n = len;
        j.scale = di ? 1.0 / (double)len : 1.0;
        j.in_role = first ? in_role : out_role;
        j.out_role = out_role;
        const int64_t idd = first ? in_dist : out_dist;
        if (inner == 1) {
            if (idd == total && out_dist == total) {
                j.count = outer * batch; j.slices = 1;
                j.in_cd = len; j.in_nd = 1; j.in_bd = 0;
                j.out_cd = len; j.out_nd = 1; j.out_bd = 0;
            } else {
                j.count = outer; j.slices = batch;
                j.in_cd = len; j.in_nd = 1; j.in_bd = idd;
                j.out_cd = len; j.out_nd = 1; j.out_bd = out_dist;
            }
            int rc = emit_axis(p, di, j, algo);
            if (rc) return rc;
        } else {
            // columns: count = inner adjacent columns, slices = outer (per batch
            // copy, or across all copies when they are packed)
            const int es = p.desc.precision ? 16 : 8;
            const bool packed = idd == total && out_dist == total;
            const int64_t nb = packed ? 1 : batch;
            for (int64_t b = 0; b < nb; ++b) {
                AxisJob k = j;
                k.count = inner;
                k.slices = packed ? outer * batch : outer;
                k.in_cd = 1; k.in_nd = inner; k.in_bd = len * inner;
                k.out_cd = 1; k.out_nd = inner; k.out_bd = len * inner;
                k.in_off = b * idd * es;
                k.out_off = b * out_dist * es;
                int rc = emit_axis(p, di, k, algo);
                if (rc) return rc;
            }
        }
        first = false;
    }
    return AFFT_OK;
}

const void* rpost_kernel(int prec, int dir);

// Real rows of n reals (h = n/2 complex points): `rows` rows, row r of slice y
// at y*bd + r*rd (complex elements of each side; the real side counts the
// row as h complex).  R2C when di == 0, C2R when di == 1.
static int emit_real_rows(Plan& p, int di, int64_t n, int64_t rows, int64_t slices, int in_role, int64_t in_rd,
                          int64_t in_bd, int out_role, int64_t out_rd, int64_t out_bd, double scale,
                          std::string& algo) {
    const int prec = p.desc.precision;
    const int es = prec ? 16 : 8;
    const int dir = di ? 1 : -1;
    const int64_t h = n / 2;
    // W_N^k, k < h
    std::vector<double> w(2 * (h + 1));
    for (int64_t k = 0; k <= h; ++k) root_of_unity(k, n, &w[2 * k], &w[2 * k + 1]);
    const void* wk;
    int rc;
    if (prec == 0) {
        std::vector<float> f(w.begin(), w.end());
        rc = upload_table(p, f.data(), f.size() * 4, &wk);
    } else {
        rc = upload_table(p, w.data(), w.size() * 8, &wk);
    }
    if (rc) return rc;
    const KernelInfo* k = (h >= 2 && is_pow2(h)) ? find_kernel(prec, (int)h, VAR_REAL, dir) : nullptr;
    if (k) {
        RArgs a{};
        KArgs tmpk{};
        if ((rc = schedule_twiddles(p, k, &tmpk.tw))) return rc;
        a.tw = tmpk.tw;
        a.wk = wk;
        a.count = rows;
        a.in_d = in_rd;
        a.out_d = out_rd;
        a.scale = di ? scale / (double)h : scale;
        for (int64_t y = 0; y < slices; ++y) {
            Step s;
            s.fn = k->fn;
            s.grid = dim3((unsigned)((rows + k->tpb - 1) / k->tpb), 1, 1);
            s.block = dim3((unsigned)k->threads, 1, 1);
            s.arg.resize(sizeof(RArgs));
            std::memcpy(s.arg.data(), &a, sizeof(RArgs));
            s.in_field = (int)offsetof(RArgs, in);
            s.out_field = (int)offsetof(RArgs, out);
            s.in_role = in_role;
            s.out_role = out_role;
            s.in_off = y * in_bd * es;
            s.out_off = y * out_bd * es;
            s.name = di ? "c2r" : "r2c";
            p.steps[di].push_back(s);
        }
        if (!algo.empty()) algo += "+";
        algo += (di ? "c2r" : "r2c") + std::to_string(n);
        return AFFT_OK;
    }
    // unfused: h-point complex transform as its own steps plus a recombination
    const size_t need = (size_t)(slices * rows * (h + 1) * es);
    if (need > p.tmp_bytes[ROLE_WORK2 - 2]) p.tmp_bytes[ROLE_WORK2 - 2] = need;
    auto post = [&](int irole, int64_t ird, int64_t ibd, int orole, int64_t ord, int64_t obd, double sc) {
        RPArgs r{};
        r.wk = wk;
        r.rows = rows;
        r.h = h;
        r.in_d = ird;
        r.out_d = ord;
        r.scale = sc;
        const int64_t per = di ? h : h + 1;
        for (int64_t y = 0; y < slices; ++y) {
            Step s;
            s.fn = rpost_kernel(prec, dir);
            s.grid = dim3((unsigned)((rows * per + 255) / 256), 1, 1);
            s.block = dim3(256, 1, 1);
            s.arg.resize(sizeof(RPArgs));
            std::memcpy(s.arg.data(), &r, sizeof(RPArgs));
            s.in_field = (int)offsetof(RPArgs, in);
            s.out_field = (int)offsetof(RPArgs, out);
            s.in_role = irole;
            s.out_role = orole;
            s.in_off = y * ibd * es;
            s.out_off = y * obd * es;
            s.name = "rpost";
            p.steps[di].push_back(s);
        }
    };
    AxisJob j{};
    j.n = h;
    j.count = rows;
    j.slices = slices;
    j.in_nd = j.out_nd = 1;
    if (!di) {
        j.in_role = in_role; j.in_cd = in_rd; j.in_bd = in_bd;
        j.out_role = ROLE_WORK2; j.out_cd = h; j.out_bd = rows * h;
        j.scale = 1.0;
        if ((rc = emit_axis(p, di, j, algo))) return rc;
        post(ROLE_WORK2, h, rows * h, out_role, out_rd, out_bd, scale);
    } else {
        post(in_role, in_rd, in_bd, ROLE_WORK2, h, rows * h, 1.0);
        j.in_role = ROLE_WORK2; j.in_cd = h; j.in_bd = rows * h;
        j.out_role = out_role; j.out_cd = out_rd; j.out_bd = out_bd;
        j.scale = scale / (double)h;
        if ((rc = emit_axis(p, di, j, algo))) return rc;
    }
    algo += di ? "(c2r)" : "(r2c)";
    return AFFT_OK;
}

// Planner's choice for AFFT_STRATEGY_AUTO on a 1-D axis: wisdom for this
// size first, then the context's cost model, else the built-in routing
// (cf. planner.go:43-78: registry, wisdom, heuristic).
// MFMA plans keep the MFMA routing: wisdom and the cost model rank the
// plain kernels, and their picks (mixed radix, Bluestein, other four-step
// splits) have no MFMA form.
uint64_t wisdom_features(const Context& c, const afft_plan_desc& d);   // tune.cpp

void choose_strategy(Plan& p, AxisJob& j) {
    if (j.in_nd != 1 || j.out_nd != 1) return;
    if (p.desc.flags & AFFT_FLAG_MFMA) return;
    Context* c = p.ctx;
    std::lock_guard<std::mutex> lock(c->mu);
    auto it = c->wisdom.find(std::make_tuple(j.n, p.desc.precision, wisdom_features(*c, p.desc)));
    Candidate pick{0, 0};
    if (it != c->wisdom.end()) {
        if (parse_strategy(it->second.algorithm, &pick)) {
            j.strategy = pick.strategy;
            j.split = pick.split;
        }
        return;
    }
    if (!c->model) return;
    std::vector<Candidate> cand = plan_candidates(p.desc);
    if (cand.size() < 2) return;
    std::vector<float> x(cand.size() * NFEAT), y(cand.size());
    for (size_t i = 0; i < cand.size(); ++i) plan_features(p.desc, cand[i], &x[i * NFEAT]);
    if (mlp_predict_host(c->model, x.data(), (int64_t)cand.size(), y.data()) != AFFT_OK) return;
    size_t best = 0;
    for (size_t i = 1; i < cand.size(); ++i)
        if (y[i] < y[best]) best = i;
    j.strategy = cand[best].strategy;
    j.split = cand[best].split;
}

static int build_plan_steps(Plan& p) {
    afft_plan_desc& d = p.desc;
    if (d.precision != AFFT_COMPLEX64 && d.precision != AFFT_COMPLEX128)
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "unknown precision");
    if (d.rank < 1 || d.rank > 8) return set_error(AFFT_ERR_INVALID_LENGTH, "rank must be 1..8");
    int64_t n = 1;
    for (int i = 0; i < d.rank; ++i) {
        if (d.dims[i] < 1) return set_error(AFFT_ERR_INVALID_LENGTH, "dimension must be >= 1");
        n *= d.dims[i];
    }
    if (d.batch == 0) d.batch = 1;
    if (d.batch < 1) return set_error(AFFT_ERR_INVALID_LENGTH, "batch must be >= 1");
    if (d.istride < 0 || d.ostride < 0 || d.idist < 0 || d.odist < 0)
        return set_error(AFFT_ERR_INVALID_STRIDE, "negative stride or distance");
    p.n = n;
    const int es = d.precision ? 16 : 8;
    const int rs = es / 2;
    const bool real = d.kind == AFFT_R2C || d.kind == AFFT_C2R;
    if (!real && d.kind != AFFT_C2C) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "unknown transform kind");
    double fscale = 1.0;
    if (d.flags & AFFT_FLAG_NORMALIZE) fscale = 1.0 / (double)n;
    if (d.flags & AFFT_FLAG_UNITARY) fscale = 1.0 / std::sqrt((double)n);

    if (!real) {
        const int64_t is = d.istride ? d.istride : 1, os = d.ostride ? d.ostride : 1;
        const int64_t id = d.idist ? d.idist : n * is, od = d.odist ? d.odist : n * os;
        if (d.rank > 1 && (is != 1 || os != 1))
            return set_error(AFFT_ERR_INVALID_STRIDE, "N-D plans take unit element stride");
        p.in_elems[0] = p.in_elems[1] = (d.batch - 1) * id + (n - 1) * is + 1;
        p.out_elems[0] = p.out_elems[1] = (d.batch - 1) * od + (n - 1) * os + 1;
        p.in_esize[0] = p.in_esize[1] = p.out_esize[0] = p.out_esize[1] = es;
        for (int di = 0; di < 2; ++di) {
            std::string algo;
            if (d.rank == 1) {
                AxisJob j{};
                j.n = n;
                j.scale = di ? 1.0 / (double)n : fscale;
                j.in_role = ROLE_SRC;
                j.out_role = ROLE_DST;
                j.count = d.batch;
                j.slices = 1;
                j.in_cd = id; j.in_nd = is;
                j.out_cd = od; j.out_nd = os;
                j.strategy = d.strategy;
                j.split = d.split;
                if (!j.strategy) choose_strategy(p, j);
                int rc = emit_axis(p, di, j, algo);
                if (rc) return rc;
            } else {
                std::vector<int> axes;
                for (int ax = d.rank - 1; ax >= 0; --ax) axes.push_back(ax);
                if (n == 1) axes = {d.rank - 1};
                int rc = emit_nd_axes(p, di, d.rank, d.dims, axes, d.batch, ROLE_SRC, id, ROLE_DST, od, algo);
                if (rc) return rc;
                if (p.steps[di].empty()) {
                    // all axes of length 1 but src != dst: plain copy
                    AxisJob j{};
                    j.n = 1; j.count = d.batch * n; j.slices = 1;
                    j.in_cd = 1; j.out_cd = 1; j.in_nd = j.out_nd = 1;
                    j.in_role = ROLE_SRC; j.out_role = ROLE_DST; j.scale = 1.0;
                    if ((rc = emit_axis(p, di, j, algo))) return rc;
                }
                if (!di && fscale != 1.0) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "normalised N-D forward");
            }
            if (di == 0) p.algo = "hip_" + algo;
        }
        return AFFT_OK;
    }

    // ---- real transforms (PlanRealT, PlanReal2D/3D) --------------------------
    const int64_t w = d.dims[d.rank - 1];
    if (w < 2 || (w & 1)) return set_error(AFFT_ERR_INVALID_LENGTH, "real transforms need an even last dimension");
    if ((d.istride && d.istride != 1) || (d.ostride && d.ostride != 1))
        return set_error(AFFT_ERR_INVALID_STRIDE, "real transforms take unit element stride");
    const int64_t h = w / 2, hw = h + 1;
    const int64_t outer = n / w;
    int64_t cd[8];
    for (int i = 0; i < d.rank; ++i) cd[i] = d.dims[i];
    cd[d.rank - 1] = hw;
    const int64_t creal = n, ccx = outer * hw;
    const int64_t rdist = d.idist ? d.idist : creal;   // real side, reals
    const int64_t cdist = d.odist ? d.odist : ccx;     // complex side, complex elements
    if (rdist & 1) return set_error(AFFT_ERR_INVALID_STRIDE, "real-side batch distance must be even");
    // [0] forward: reals in, complex out; [1] inverse: complex in, reals out
    p.in_elems[0] = (d.batch - 1) * rdist + creal;
    p.out_elems[0] = (d.batch - 1) * cdist + ccx;
    p.in_elems[1] = p.out_elems[0];
    p.out_elems[1] = p.in_elems[0];
    p.in_esize[0] = p.out_esize[1] = rs;
    p.out_esize[0] = p.in_esize[1] = es;
    const int64_t rd2 = rdist / 2;   // real batch distance in complex units
    for (int di = 0; di < 2; ++di) {
        std::string algo;
        std::vector<int> axes;
        for (int ax = d.rank - 2; ax >= 0; --ax) axes.push_back(ax);
        int rc;
        // rows are contiguous within one batch copy: outer rows of h / hw
        const bool flat = (rdist == creal && cdist == ccx);
        const int64_t rows = flat ? outer * d.batch : outer;
        const int64_t slices = flat ? 1 : d.batch;
        if (!di) {
            rc = emit_real_rows(p, 0, w, rows, slices, ROLE_SRC, h, rd2, ROLE_DST, hw, cdist, fscale, algo);
            if (rc) return rc;
            if (!axes.empty()) {
                if (fscale != 1.0) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "normalised N-D real forward");
                rc = emit_nd_axes(p, 0, d.rank, cd, axes, d.batch, ROLE_DST, cdist, ROLE_DST, cdist, algo);
                if (rc) return rc;
            }
            p.algo = "hip_" + algo;
        } else {
            int src_role = ROLE_SRC;
            int64_t src_dist = cdist;
            if (!axes.empty()) {
                // outer axes first (inverse order of the forward), into WORK
                std::vector<int> inv_axes(axes.rbegin(), axes.rend());
                const size_t need = (size_t)(d.batch * ccx * es);
                if (need > p.tmp_bytes[ROLE_WORK - 2]) p.tmp_bytes[ROLE_WORK - 2] = need;
                rc = emit_nd_axes(p, 1, d.rank, cd, inv_axes, d.batch, ROLE_SRC, cdist, ROLE_WORK, ccx, algo);
                if (rc) return rc;
                src_role = ROLE_WORK;
                src_dist = ccx;
            }
            const bool flat2 = (rdist == creal && src_dist == ccx);
            const int64_t rows2 = flat2 ? outer * d.batch : outer;
            const int64_t slices2 = flat2 ? 1 : d.batch;
            rc = emit_real_rows(p, 1, w, rows2, slices2, src_role, hw, src_dist, ROLE_DST, h, rd2, 1.0, algo);
            if (rc) return rc;
        }
    }
    return AFFT_OK;
}

// Enqueues the steps of direction di on s, with the roles resolved to src,
// dst and the workspace buffers tmp[0 .. NTMP).
int launch_steps(Plan& p, int di, void* dst, const void* src, void* const* tmp, hipStream_t s) {
    void* base[ROLE_COUNT] = {const_cast<void*>(src), dst, tmp[0], tmp[1], tmp[2], tmp[3]};
    for (Step& st : p.steps[di]) {
        char* ip = (char*)base[st.in_role] + st.in_off;
        char* op = (char*)base[st.out_role] + st.out_off;
        std::memcpy(st.arg.data() + st.in_field, &ip, sizeof(void*));
        std::memcpy(st.arg.data() + st.out_field, &op, sizeof(void*));
        void* args[] = {st.arg.data()};
        AFFT_HIP(hipLaunchKernel(st.fn, st.grid, st.block, args, (size_t)st.lds_bytes, s));
    }
    return AFFT_OK;
}

// A plan's workspace (tmp[]) is shared by all its executions.  When one
// execution is enqueued on a different stream than the previous one, the new
// stream first waits for the previous use to finish, so a plan may be driven
// from several streams (and host threads) without racing on its scratch.
// Under stream capture the caller orders the graph's nodes and no event is
// touched.
int exec_plan(Plan& p, int dir, void* dst, const void* src, hipStream_t s) {
    const int di = dir == AFFT_INVERSE ? 1 : 0;
    bool ws = false;
    for (int t = 0; t < NTMP; ++t) {
        if (!p.tmp_bytes[t]) continue;
        ws = true;
        if (!p.tmp[t]) AFFT_HIP(hipMalloc(&p.tmp[t], p.tmp_bytes[t]));
    }
    if (ws) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        AFFT_HIP(hipStreamIsCapturing(s, &cs));
        if (cs != hipStreamCaptureStatusNone) {
            ws = false;
        } else {
            if (!p.ws_done) AFFT_HIP(hipEventCreateWithFlags(&p.ws_done, hipEventDisableTiming));
            if (p.ws_used && p.ws_stream != s) AFFT_HIP(hipStreamWaitEvent(s, p.ws_done, 0));
        }
    }
    int rc = launch_steps(p, di, dst, src, p.tmp, s);
    if (rc) return rc;
    if (ws) {
        AFFT_HIP(hipEventRecord(p.ws_done, s));
        p.ws_stream = s;
        p.ws_used = true;
    }
    return AFFT_OK;
}

// Launches that write the caller's output store it non-temporally when that
// output is at least kStreamOutBytes: larger outputs would be evicted from the
// last-level cache before anyone re-reads them, and bypassing it measured
// about 2 % faster on the 512 MiB batches.  A launch that also reads the
// caller's input of that size streams its loads as well (one pass from
// source to destination): about 3 % more on the N=4096 batch, while
// streaming loads alone lose.
constexpr size_t kStreamOutBytes = size_t(128) << 20;

int build_plan(Plan& p) {
    int rc = build_plan_steps(p);
    if (rc) return rc;
    for (int di = 0; di < 2; ++di) {
        const size_t ob = (size_t)p.out_elems[di] * p.out_esize[di];
        const size_t ib = (size_t)p.in_elems[di] * p.in_esize[di];
        if (ob < kStreamOutBytes) continue;
        for (Step& st : p.steps[di]) {
            if (st.out_role != ROLE_DST || st.nt_field < 0) continue;
            int flags = NT_STORE;
            if (st.in_role == ROLE_SRC && ib >= kStreamOutBytes) flags |= NT_LOAD;
            std::memcpy(st.arg.data() + st.nt_field, &flags, sizeof(int));
        }
    }
    return AFFT_OK;
}

}  // namespace afft
