// Full-spectrum forms of the real N-D plans: PlanReal2D/3D.ForwardFull and
// InverseFull (plan_real_2d.go:231-276, :370-389; plan_real_3d.go:241-285,
// :392-419).
//
// ForwardFull runs the plan's R2C into a compact half spectrum (rows of
// W/2+1 bins) and expands it: bin (i, c) of the full D x ... x W spectrum is
// the half-spectrum bin for c <= W/2 and conj(X[-i mod dims, W - c]) above,
// the conjugate symmetry of a real input's DFT.  InverseFull keeps the first
// W/2+1 bins of every full row and runs the plan's C2R on them, as the
// reference does.  Both take a single transform (batch 1), as the reference's
// length checks do.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "afft_internal.hpp"
#include "fft_common.hpp"

namespace afft {

struct FullArgs {
    const void* src;
    void* dst;
    int w;              // last dimension (real points per row)
    int hw;             // w/2 + 1 bins per half-spectrum row
    int nouter;         // number of outer dimensions
    int dims[8];        // outer dimensions, outermost first
};

// One block per full-spectrum row r (outer index tuple in row-major order).
template <typename T>
__global__ void __launch_bounds__(256) k_expand_full(FullArgs a) {
    const long long r = blockIdx.x;
    // the mirrored row: every outer index negated modulo its dimension
    long long rem = r, mr = 0, scale = 1;
    for (int k = a.nouter - 1; k >= 0; --k) {
        const int d = a.dims[k];
        const int i = (int)(rem % d);
        rem /= d;
        mr += (long long)((d - i) % d) * scale;
        scale *= d;
    }
    const cx<T>* __restrict__ h = (const cx<T>*)a.src + r * a.hw;
    const cx<T>* __restrict__ hm = (const cx<T>*)a.src + mr * a.hw;
    cx<T>* __restrict__ f = (cx<T>*)a.dst + r * a.w;
    for (int c = threadIdx.x; c < a.w; c += blockDim.x) {
        if (c < a.hw) {
            f[c] = h[c];
        } else {
            const cx<T> v = hm[a.w - c];
            f[c] = mk<T>(v.x, -v.y);
        }
    }
}

// The first hw bins of every full row, packed into half-spectrum rows.
template <typename T>
__global__ void __launch_bounds__(256) k_compact_half(FullArgs a) {
    const long long r = blockIdx.x;
    const cx<T>* __restrict__ f = (const cx<T>*)a.src + r * a.w;
    cx<T>* __restrict__ h = (cx<T>*)a.dst + r * a.hw;
    for (int c = threadIdx.x; c < a.hw; c += blockDim.x) h[c] = f[c];
}

static int full_check(afft_plan p) {
    const afft_plan_desc& d = p->desc;
    if (d.kind != AFFT_R2C && d.kind != AFFT_C2R)
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "full-spectrum transforms need a real plan");
    if (d.batch > 1) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "full-spectrum transforms take one transform");
    if ((d.istride > 1) || (d.ostride > 1))
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "full-spectrum transforms need the contiguous layout");
    if (d.rank < 1 || d.rank > 8 || d.dims[d.rank - 1] > (1 << 30))
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "full-spectrum row too long");
    if (p->steps[0].empty() || p->steps[1].empty())
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "full-spectrum transforms need both directions of the plan");
    return AFFT_OK;
}

// Enqueues a full-spectrum transform on s; the caller holds the plan lock.
static int exec_full(afft_plan p, int dir, void* dst, const void* src, hipStream_t s) {
    const afft_plan_desc& d = p->desc;
    const int es = d.precision ? 16 : 8;
    FullArgs a{};
    a.w = (int)d.dims[d.rank - 1];
    a.hw = a.w / 2 + 1;
    a.nouter = d.rank - 1;
    long long rows = 1;
    for (int k = 0; k < a.nouter; ++k) {
        a.dims[k] = (int)d.dims[k];
        rows *= d.dims[k];
    }
    if (rows > 0x7fffffffLL) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "too many full-spectrum rows");
    const size_t half_bytes = (size_t)rows * a.hw * es;
    if (p->full_bytes < half_bytes) {
        if (p->full) AFFT_HIP(hipFree(p->full));
        p->full = nullptr;
        p->full_bytes = 0;
        AFFT_HIP(hipMalloc(&p->full, half_bytes));
        p->full_bytes = half_bytes;
    }
    // The half-spectrum buffer is shared by the plan's executions: order
    // them across streams as exec_plan does for its workspace.
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    AFFT_HIP(hipStreamIsCapturing(s, &cs));
    const bool order = cs == hipStreamCaptureStatusNone;
    if (order) {
        if (!p->ws_done) AFFT_HIP(hipEventCreateWithFlags(&p->ws_done, hipEventDisableTiming));
        if (p->ws_used && p->ws_stream != s) AFFT_HIP(hipStreamWaitEvent(s, p->ws_done, 0));
    }
    const void* fn;
    int rc;
    if (dir == AFFT_FORWARD) {
        if ((rc = exec_plan(*p, AFFT_FORWARD, p->full, src, s))) return rc;
        a.src = p->full;
        a.dst = dst;
        fn = d.precision ? (const void*)&k_expand_full<double> : (const void*)&k_expand_full<float>;
        void* args[] = {&a};
        AFFT_HIP(hipLaunchKernel(fn, dim3((unsigned)rows), dim3(256), args, 0, s));
    } else {
        a.src = src;
        a.dst = p->full;
        fn = d.precision ? (const void*)&k_compact_half<double> : (const void*)&k_compact_half<float>;
        void* args[] = {&a};
        AFFT_HIP(hipLaunchKernel(fn, dim3((unsigned)rows), dim3(256), args, 0, s));
        if ((rc = exec_plan(*p, AFFT_INVERSE, dst, p->full, s))) return rc;
    }
    if (order) {
        AFFT_HIP(hipEventRecord(p->ws_done, s));
        p->ws_stream = s;
        p->ws_used = true;
    }
    return AFFT_OK;
}

}  // namespace afft

using namespace afft;

extern "C" {

int afft_plan_exec_full_device(afft_plan p, int direction, void* dst, const void* src, void* stream) {
    if (!p) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "nil plan");
    if (!dst || !src) return set_error(AFFT_ERR_NIL_SLICE, "dst or src is nil");
    if (direction != AFFT_FORWARD && direction != AFFT_INVERSE)
        return set_error(AFFT_ERR_INVALID_ARGUMENT, "direction must be -1 or +1");
    int rc = full_check(p);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(p->mu);
    AFFT_HIP(hipSetDevice(p->ctx->device));
    rc = exec_full(p, direction, dst, src, stream ? (hipStream_t)stream : p->stream);
    if (rc == AFFT_OK) clear_error();
    return rc;
}

// Host slices: the reference's checks in its order (nil, then exact lengths:
// n reals and n complex bins), then staging through device memory.
static int host_full(afft_plan p, int dir, void* dst, int64_t dst_len, const void* src, int64_t src_len) {
    if (!p) return set_error(AFFT_ERR_NOT_IMPLEMENTED, "nil plan");
    if (!dst || !src) return set_error(AFFT_ERR_NIL_SLICE, "dst or src is nil");
    int rc = full_check(p);
    if (rc) return rc;
    if (src_len != p->n || dst_len != p->n)
        return set_error(AFFT_ERR_LENGTH_MISMATCH, "full-spectrum slices must hold exactly the plan's points");
    const size_t rs = p->desc.precision ? 8 : 4, cs = 2 * rs;
    const size_t ib = (size_t)p->n * (dir == AFFT_FORWARD ? rs : cs);
    const size_t ob = (size_t)p->n * (dir == AFFT_FORWARD ? cs : rs);
    std::lock_guard<std::mutex> lock(p->mu);
    AFFT_HIP(hipSetDevice(p->ctx->device));
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
    if ((rc = exec_full(p, dir, p->stage_out, p->stage_in, p->stream))) return rc;
    AFFT_HIP(hipMemcpyAsync(dst, p->stage_out, ob, hipMemcpyDeviceToHost, p->stream));
    AFFT_HIP(hipStreamSynchronize(p->stream));
    clear_error();
    return AFFT_OK;
}

int afft_plan_forward_full(afft_plan p, void* dst, int64_t dst_len, const void* src, int64_t src_len) {
    return host_full(p, AFFT_FORWARD, dst, dst_len, src, src_len);
}

int afft_plan_inverse_full(afft_plan p, void* dst, int64_t dst_len, const void* src, int64_t src_len) {
    return host_full(p, AFFT_INVERSE, dst, dst_len, src, src_len);
}

}  // extern "C"
