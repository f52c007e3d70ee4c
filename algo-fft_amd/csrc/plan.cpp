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
#include <cstring>

#include "afft_internal.hpp"

namespace afft {

Plan::~Plan() {
    if (ctx) (void)hipSetDevice(ctx->device);
    for (void* q : owned) (void)hipFree(q);
    for (void*& t : tmp)
        if (t) (void)hipFree(t);
    if (stage_in) (void)hipFree(stage_in);
    if (stage_out) (void)hipFree(stage_out);
    if (stream) (void)hipStreamDestroy(stream);
}

static bool is_pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }
static int ilog2(int64_t n) {
    int l = 0;
    while ((1LL << l) < n) ++l;
    return l;
}

// Upload a host table, owned by the plan.
static int upload_table(Plan& p, const void* host, size_t bytes, const void** dev) {
    void* d = nullptr;
    AFFT_HIP(hipMalloc(&d, bytes ? bytes : 16));
    p.owned.push_back(d);
    if (bytes) AFFT_HIP(hipMemcpy(d, host, bytes, hipMemcpyHostToDevice));
    *dev = d;
    return AFFT_OK;
}

// Twiddle table of a compile-time schedule: for pass i >= 1 with
// Ns = rad[0]*...*rad[i-1] and radix R: entries W_{Ns*R}^{k*r}, k < Ns,
// 1 <= r < R, stored k-major.  Generated in FP64 and rounded once.
static int schedule_twiddles(Plan& p, const KernelInfo* k, const void** dev) {
    std::vector<double> tw;
    int64_t ns = k->rad[0];
    for (int i = 1; i < k->nrad; ++i) {
        const int R = k->rad[i];
        for (int64_t kk = 0; kk < ns; ++kk)
            for (int r = 1; r < R; ++r) {
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
static int split_twiddles(Plan& p, int prec, int64_t L, int* shift, const void** lo, const void** hi) {
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
};

static int emit_axis(Plan& p, int di, const AxisJob& j, std::string& algo);

static int emit_single(Plan& p, int di, const AxisJob& j, std::string& algo) {
    const int prec = p.desc.precision;
    const int dir = di ? 1 : -1;
    const int es = prec ? 16 : 8;
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
static int emit_fourstep(Plan& p, int di, const AxisJob& j, std::string& algo) {
    const int prec = p.desc.precision;
    const int dir = di ? 1 : -1;
    const int es = prec ? 16 : 8;
    if (j.in_nd != 1 || j.out_nd != 1 || !is_pow2(j.n))
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "four-step needs a unit-stride power-of-two axis");
    const int lg = ilog2(j.n);
    const int64_t n1 = 1LL << (lg / 2), n2 = j.n / n1;
    if (n2 > max_single_pass(prec)) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "transform too large");
    const KernelInfo* ka = find_kernel(prec, (int)n1, VAR_4A, dir);
    const KernelInfo* kb = find_kernel(prec, (int)n2, VAR_4B, dir);
    if (!ka || !kb) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "missing four-step kernel");
    // every (c, y) transform is one slice of the four-step passes
    const int64_t total = j.count * j.slices;
    const int64_t need = total * j.n * es;
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
    // the transforms (c, y): c-major inside each y slice
    for (int64_t y = 0; y < j.slices; ++y) {
        // transforms of this slice are contiguous in c with distance cd; a
        // launch covers all c with blockIdx.y = c
        KArgs aa = a, bb = b;
        aa.in_bd = j.in_cd;
        bb.out_bd = j.out_cd;
        Step sa = make_step(ka, aa, (n2 + ka->tpb - 1) / ka->tpb, 1, "4stepA");
        sa.in_role = j.in_role;
        sa.in_off = j.in_off + y * j.in_bd * es;
        sa.out_role = ROLE_TMP0;
        sa.out_off = y * j.count * j.n * es;
        push_sliced(p.steps[di], sa, j.count, j.in_cd * es, j.n * es);
        Step sb = make_step(kb, bb, (n1 + kb->tpb - 1) / kb->tpb, 1, "4stepB");
        sb.in_role = ROLE_TMP0;
        sb.in_off = y * j.count * j.n * es;
        sb.out_role = j.out_role;
        sb.out_off = j.out_off + y * j.out_bd * es;
        push_sliced(p.steps[di], sb, j.count, j.n * es, j.out_cd * es);
    }
    if (!algo.empty()) algo += "+";
    algo += "fourstep" + std::to_string(n1) + "x" + std::to_string(n2);
    return AFFT_OK;
}

static int emit_axis(Plan& p, int di, const AxisJob& j, std::string& algo) {
    if (j.n == 1) {
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "length-1 axis");
    }
    if (is_pow2(j.n) && j.n <= max_single_pass(p.desc.precision)) return emit_single(p, di, j, algo);
    if (is_pow2(j.n)) return emit_fourstep(p, di, j, algo);
    return set_error(AFFT_ERR_NOT_IMPLEMENTED, "size " + std::to_string(j.n) + " not supported yet");
}

int build_plan(Plan& p) {
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
    if (d.kind != AFFT_C2C) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "real transforms not built yet");

    const int64_t is = d.istride ? d.istride : 1, os = d.ostride ? d.ostride : 1;
    const int64_t id = d.idist ? d.idist : n * is, od = d.odist ? d.odist : n * os;
    if (d.rank > 1 && (is != 1 || os != 1))
        return set_error(AFFT_ERR_INVALID_STRIDE, "N-D plans take unit element stride");
    // extent of the data each side addresses (elements)
    p.in_elems[0] = p.in_elems[1] = (d.batch - 1) * id + (n - 1) * is + 1;
    p.out_elems[0] = p.out_elems[1] = (d.batch - 1) * od + (n - 1) * os + 1;
    p.in_esize[0] = p.in_esize[1] = p.out_esize[0] = p.out_esize[1] = es;

    for (int di = 0; di < 2; ++di) {
        std::string algo;
        // axes last to first (plan_nd.go:405); the last axis reads src, the
        // others work in place on dst
        int64_t inner = 1;
        for (int ax = d.rank - 1; ax >= 0; --ax) {
            const int64_t len = d.dims[ax];
            const bool first = ax == d.rank - 1;
            if (len == 1) {
                if (first && d.rank == 1) {
                    // a 1-point transform is a copy (scaled by 1 either way)
                    AxisJob j{};
                    (void)j;
                }
                inner *= len;
                continue;
            }
            AxisJob j{};
            j.n = len;
            j.scale = di ? 1.0 / (double)len : 1.0;
            j.in_role = first ? ROLE_SRC : ROLE_DST;
            j.out_role = ROLE_DST;
            if (d.rank == 1) {
                j.count = d.batch;
                j.slices = 1;
                j.in_cd = id; j.in_nd = is; j.in_bd = 0;
                j.out_cd = od; j.out_nd = os; j.out_bd = 0;
            } else {
                // axis with stride `inner`; outer = everything before it, per batch
                const int64_t outer = (n / (len * inner));
                const bool src_side = first;
                const int64_t bdist_in = src_side ? id : od;
                (void)bdist_in;
                if (inner == 1) {
                    // rows: count = outer*batch when distances are packed
                    if ((src_side ? id : od) == n && od == n) {
                        j.count = outer * d.batch;
                        j.slices = 1;
                        j.in_cd = len; j.in_nd = 1; j.in_bd = 0;
                        j.out_cd = len; j.out_nd = 1; j.out_bd = 0;
                    } else {
                        j.count = outer;
                        j.slices = d.batch;
                        j.in_cd = len; j.in_nd = 1; j.in_bd = src_side ? id : od;
                        j.out_cd = len; j.out_nd = 1; j.out_bd = od;
                    }
                } else {
                    j.count = inner;
                    j.slices = outer * d.batch;
                    if ((src_side ? id : od) != n || od != n)
                        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "batched N-D with padded distance");
                    j.in_cd = 1; j.in_nd = inner; j.in_bd = len * inner;
                    j.out_cd = 1; j.out_nd = inner; j.out_bd = len * inner;
                }
            }
            int rc = emit_axis(p, di, j, algo);
            if (rc) return rc;
            inner *= len;
        }
        if (p.steps[di].empty()) {
            // every axis has length 1: a plain copy of each point
            return set_error(AFFT_ERR_NOT_IMPLEMENTED, "all axes have length 1");
        }
        if (di == 0) p.algo = "hip_" + algo;
    }
    return AFFT_OK;
}

int exec_plan(Plan& p, int dir, void* dst, const void* src, hipStream_t s) {
    const int di = dir == AFFT_INVERSE ? 1 : 0;
    for (int t = 0; t < 2; ++t) {
        if (p.tmp_bytes[t] && !p.tmp[t]) AFFT_HIP(hipMalloc(&p.tmp[t], p.tmp_bytes[t]));
    }
    void* base[ROLE_COUNT] = {const_cast<void*>(src), dst, p.tmp[0], p.tmp[1]};
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

}  // namespace afft
