// Convolution and correlation on the device: the first callers of the
// transform path (Convolve, convolve.go:9-115; ConvolveReal,
// convolve_real.go:9; CrossCorrelate / AutoCorrelate, correlate.go:6-79).
//
// Linear convolution of a (length na) and b (length nb), L = na + nb - 1:
// both inputs are zero-padded to an FFT length F >= L, transformed, multiplied
// pointwise, transformed back and cropped to L.  F is the smallest 7-smooth
// length >= L (the reference uses F = L for complex inputs and the next power
// of two for real ones; any F >= L gives the same linear convolution).
// Cross-correlation is the convolution with b reversed and conjugated, which
// the padding kernel does on the fly.  Real convolution (ConvolveReal) pads
// real rows to an even 7-smooth F, runs R2C on both, multiplies the half
// spectra and returns through C2R.
#include <algorithm>
#include <cstring>

#include "afft_internal.hpp"

namespace afft {

struct PadArgs {
    const void* src;
    void* dst;
    long long n, f;         // source length, padded length
    long long sdist, ddist; // per-batch distances (elements)
    long long batch;
    int revconj;            // 1: dst[i] = conj(src[n-1-i])
};

template <typename T>
__global__ void __launch_bounds__(256) k_pad(PadArgs a) {
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= a.batch * a.f) return;
    const long long b = idx / a.f, i = idx - b * a.f;
    cx<T> v = mk<T>(0, 0);
    if (i < a.n) {
        const cx<T>* s = (const cx<T>*)a.src + b * a.sdist;
        v = a.revconj ? conjx(s[a.n - 1 - i]) : s[i];
    }
    ((cx<T>*)a.dst)[b * a.ddist + i] = v;
}

struct MulArgs {
    void* x;                // x *= y elementwise
    const void* y;
    long long count;
};

template <typename T>
__global__ void __launch_bounds__(256) k_cmul(MulArgs a) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.count) return;
    cx<T>* x = (cx<T>*)a.x;
    x[i] = cmul(x[i], ((const cx<T>*)a.y)[i]);
}

template <typename T>
__global__ void __launch_bounds__(256) k_pad_real(PadArgs a) {
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= a.batch * a.f) return;
    const long long b = idx / a.f, i = idx - b * a.f;
    ((T*)a.dst)[b * a.ddist + i] = i < a.n ? ((const T*)a.src)[b * a.sdist + i] : T(0);
}

struct CropArgs {
    const void* src;
    void* dst;
    long long len, f, ddist, batch;
};

template <typename T>
__global__ void __launch_bounds__(256) k_crop(CropArgs a) {
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= a.batch * a.len) return;
    const long long b = idx / a.len, i = idx - b * a.len;
    ((cx<T>*)a.dst)[b * a.ddist + i] = ((const cx<T>*)a.src)[b * a.f + i];
}

template <typename T>
__global__ void __launch_bounds__(256) k_crop_real(CropArgs a) {
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= a.batch * a.len) return;
    const long long b = idx / a.len, i = idx - b * a.len;
    ((T*)a.dst)[b * a.ddist + i] = ((const T*)a.src)[b * a.f + i];
}

}  // namespace afft

using namespace afft;

struct afft_conv_s {
    afft_context ctx = nullptr;
    int prec = 0;
    int mode = 0;             // AFFT_CONV_LINEAR, AFFT_CONV_CORRELATE or AFFT_CONV_REAL
    int64_t na = 0, nb = 0, len = 0, f = 0, batch = 1;
    afft_plan plan = nullptr; // forward: 2*batch transforms of f (a rows then b rows)
    afft_plan inv = nullptr;  // inverse: the batch product rows
    void* work = nullptr;     // 2 * batch * f complex (real mode: 2 * batch * f reals)
    void* spec = nullptr;     // real mode: 2 * batch * (f/2 + 1) complex spectra
    hipStream_t stream = nullptr;
    std::mutex mu;
    ~afft_conv_s() {
        if (ctx) (void)hipSetDevice(ctx->device);
        if (plan) afft_plan_destroy(plan);
        if (inv) afft_plan_destroy(inv);
        if (work) (void)hipFree(work);
        if (spec) (void)hipFree(spec);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// Smallest 7-smooth length >= n (even if `even`: the real transforms need it).
static int64_t smooth_at_least(int64_t n, bool even = false) {
    for (int64_t m = std::max<int64_t>(n, even ? 2 : 1);; ++m) {
        if (even && (m & 1)) continue;
        int64_t t = m;
        for (int q : {2, 3, 5, 7})
            while (t % q == 0) t /= q;
        if (t == 1) return m;
    }
}

static int launch(const void* fn, int64_t total, void* args, hipStream_t s) {
    void* a[] = {args};
    AFFT_HIP(hipLaunchKernel(fn, dim3((unsigned)((total + 255) / 256)), dim3(256), a, 0, s));
    return AFFT_OK;
}

extern "C" {

// Plan for `batch` convolutions (AFFT_CONV_LINEAR) or cross-correlations
// (AFFT_CONV_CORRELATE) of complex sequences, or convolutions of real ones
// (AFFT_CONV_REAL), of lengths na and nb; inputs a: batch rows of na
// elements, b: batch rows of nb; output batch rows of na + nb - 1.
int afft_conv_create(afft_context ctx, int precision, int64_t na, int64_t nb, int64_t batch, int mode,
                     afft_conv* out) {
    if (!ctx || !out) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL handle");
    *out = nullptr;
    if (na < 1 || nb < 1 || batch < 1) return set_error(AFFT_ERR_INVALID_LENGTH, "lengths must be >= 1");
    if (precision != AFFT_COMPLEX64 && precision != AFFT_COMPLEX128)
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "unknown precision");
    if (mode != AFFT_CONV_LINEAR && mode != AFFT_CONV_CORRELATE && mode != AFFT_CONV_REAL)
        return set_error(AFFT_ERR_INVALID_ARGUMENT, "unknown convolution mode");
    AFFT_HIP(hipSetDevice(ctx->device));
    auto* c = new afft_conv_s();
    c->ctx = ctx;
    c->prec = precision;
    c->mode = mode;
    c->na = na;
    c->nb = nb;
    c->batch = batch;
    c->len = na + nb - 1;
    const bool real = mode == AFFT_CONV_REAL;
    c->f = smooth_at_least(c->len, real);
    afft_plan_desc d{};
    d.kind = real ? AFFT_R2C : AFFT_C2C;
    d.precision = precision;
    d.rank = 1;
    d.dims[0] = c->f;
    d.batch = 2 * batch;
    int rc = afft_plan_create_desc(ctx, &d, &c->plan);
    d.batch = batch;
    if (!rc) rc = afft_plan_create_desc(ctx, &d, &c->inv);
    hipError_t e = hipSuccess;
    const size_t ce = precision ? 16 : 8;
    if (!rc) e = hipMalloc(&c->work, (size_t)(2 * batch * c->f) * (real ? ce / 2 : ce));
    if (!rc && e == hipSuccess && real) e = hipMalloc(&c->spec, (size_t)(2 * batch * (c->f / 2 + 1)) * ce);
    if (!rc && e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (!rc && e != hipSuccess) rc = set_error(AFFT_ERR_DEVICE, std::string("conv setup: ") + hipGetErrorString(e));
    if (rc) {
        std::string keep = afft_last_error();
        delete c;
        set_error(rc, keep);
        return rc;
    }
    *out = c;
    clear_error();
    return AFFT_OK;
}

// Device pointers; enqueued on `stream` (NULL: the plan's own).
int afft_conv_exec(afft_conv c, void* dst, const void* a, const void* b, void* stream) {
    if (!c) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL handle");
    if (!dst || !a || !b) return set_error(AFFT_ERR_NIL_SLICE, "nil slice");
    std::lock_guard<std::mutex> l(c->mu);
    AFFT_HIP(hipSetDevice(c->ctx->device));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    if (c->mode == AFFT_CONV_REAL) {
        // real rows padded to f, half spectra of f/2 + 1, product, C2R, crop
        const int rs = c->prec ? 8 : 4, cs = 2 * rs;
        const int64_t h = c->f / 2 + 1;
        char* ra = (char*)c->work;
        char* rb = ra + (size_t)c->batch * c->f * rs;
        char* sa = (char*)c->spec;
        char* sb = sa + (size_t)c->batch * h * cs;
        const void* fpad = c->prec ? (const void*)&k_pad_real<double> : (const void*)&k_pad_real<float>;
        const void* fmul = c->prec ? (const void*)&k_cmul<double> : (const void*)&k_cmul<float>;
        const void* fcrop = c->prec ? (const void*)&k_crop_real<double> : (const void*)&k_crop_real<float>;
        PadArgs pa{a, ra, c->na, c->f, c->na, c->f, c->batch, 0};
        int rc = launch(fpad, c->batch * c->f, &pa, s);
        PadArgs pb{b, rb, c->nb, c->f, c->nb, c->f, c->batch, 0};
        if (!rc) rc = launch(fpad, c->batch * c->f, &pb, s);
        if (!rc) rc = afft_plan_exec_device(c->plan, AFFT_FORWARD, c->spec, c->work, s);
        MulArgs m{sa, sb, c->batch * h};
        if (!rc) rc = launch(fmul, m.count, &m, s);
        if (!rc) rc = afft_plan_exec_device(c->inv, AFFT_INVERSE, ra, sa, s);
        CropArgs cr{ra, dst, c->len, c->f, c->len, c->batch};
        if (!rc) rc = launch(fcrop, c->batch * c->len, &cr, s);
        if (!rc) clear_error();
        return rc;
    }
    const int es = c->prec ? 16 : 8;
    char* wa = (char*)c->work;
    char* wb = wa + (size_t)c->batch * c->f * es;
    const void* fpad = c->prec ? (const void*)&k_pad<double> : (const void*)&k_pad<float>;
    const void* fmul = c->prec ? (const void*)&k_cmul<double> : (const void*)&k_cmul<float>;
    const void* fcrop = c->prec ? (const void*)&k_crop<double> : (const void*)&k_crop<float>;
    PadArgs pa{a, wa, c->na, c->f, c->na, c->f, c->batch, 0};
    int rc = launch(fpad, c->batch * c->f, &pa, s);
    PadArgs pb{b, wb, c->nb, c->f, c->nb, c->f, c->batch, c->mode == AFFT_CONV_CORRELATE};
    if (!rc) rc = launch(fpad, c->batch * c->f, &pb, s);
    if (!rc) rc = afft_plan_exec_device(c->plan, AFFT_FORWARD, c->work, c->work, s);
    MulArgs m{wa, wb, c->batch * c->f};
    if (!rc) rc = launch(fmul, m.count, &m, s);
    if (!rc) rc = afft_plan_exec_device(c->inv, AFFT_INVERSE, wa, wa, s);
    CropArgs cr{wa, dst, c->len, c->f, c->len, c->batch};
    if (!rc) rc = launch(fcrop, c->batch * c->len, &cr, s);
    if (!rc) clear_error();
    return rc;
}

// Host-slice convenience in the shape of Convolve / CrossCorrelate: one
// convolution, synchronised.
int afft_convolve_host(afft_context ctx, int precision, int mode, void* dst, int64_t dst_len, const void* a,
                       int64_t na, const void* b, int64_t nb) {
    if (!ctx) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL context");
    if (!dst || !a || !b) return set_error(AFFT_ERR_NIL_SLICE, "nil slice");          // convolve.go:10-12
    if (na < 1 || nb < 1) return set_error(AFFT_ERR_INVALID_LENGTH, "empty input");   // :14-16
    if (dst_len != na + nb - 1) return set_error(AFFT_ERR_LENGTH_MISMATCH, "dst must hold len(a)+len(b)-1");
    afft_conv c = nullptr;
    int rc = afft_conv_create(ctx, precision, na, nb, 1, mode, &c);
    if (rc) return rc;
    // element size of the host slices: complex, or real for AFFT_CONV_REAL
    const int es = (precision ? 16 : 8) / (mode == AFFT_CONV_REAL ? 2 : 1);
    void *da = nullptr, *db = nullptr, *dd = nullptr;
    hipError_t e = hipMalloc(&da, na * es);
    if (e == hipSuccess) e = hipMalloc(&db, nb * es);
    if (e == hipSuccess) e = hipMalloc(&dd, dst_len * es);
    if (e == hipSuccess) e = hipMemcpy(da, a, na * es, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, b, nb * es, hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = set_error(AFFT_ERR_DEVICE, std::string("convolve: ") + hipGetErrorString(e));
    if (!rc) rc = afft_conv_exec(c, dd, da, db, nullptr);
    if (!rc) {
        e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipMemcpy(dst, dd, dst_len * es, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = set_error(AFFT_ERR_DEVICE, std::string("convolve: ") + hipGetErrorString(e));
    }
    for (void* q : {da, db, dd})
        if (q) (void)hipFree(q);
    afft_conv_destroy(c);
    if (!rc) clear_error();
    return rc;
}

int afft_conv_destroy(afft_conv c) {
    delete c;
    clear_error();
    return AFFT_OK;
}

}  // extern "C"
