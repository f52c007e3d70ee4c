// Zero-validation fast paths: FastPlan (plan_fast.go:21-133) and
// Plan.Clone / Executor (plan.go:1300, executor.go:6-43).
//
// A fast plan is an ordinary plan whose forward and inverse step lists are
// one launch each; creation copies those launches (kernel, grid, block,
// argument block with the pointer fields' offsets) out of it, and every call
// afterwards only patches the two pointers into the argument block and
// launches.  No status is returned and no lock is taken, like the
// reference's FastPlan.Forward, which returns nothing either.
#include <cstring>

#include "afft_internal.hpp"

using namespace afft;

struct afft_fast_plan_s {
    afft_plan plan = nullptr;     // owns the tables and the stream
    hipStream_t stream = nullptr;
    Step step[2];
    ~afft_fast_plan_s() {
        if (plan) afft_plan_destroy(plan);
    }
};

static inline void fast_launch(afft_fast_plan fp, int di, void* dst, const void* src) {
    Step& st = fp->step[di];
    const char* ip = (const char*)src + st.in_off;
    char* op = (char*)dst + st.out_off;
    std::memcpy(st.arg.data() + st.in_field, &ip, sizeof(void*));
    std::memcpy(st.arg.data() + st.out_field, &op, sizeof(void*));
    void* args[] = {st.arg.data()};
    (void)hipLaunchKernel(st.fn, st.grid, st.block, args, (size_t)st.lds_bytes, fp->stream);
}

extern "C" {

int afft_plan_clone(afft_plan p, afft_plan* out) {
    if (!p || !out) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL handle");
    *out = nullptr;
    afft_plan_desc d = p->desc;
    return afft_plan_create_desc(static_cast<afft_context>(p->ctx), &d, out);
}

int afft_fast_plan_create(afft_context ctx, int64_t n, int64_t batch, int precision, void* stream,
                          afft_fast_plan* out) {
    if (out) *out = nullptr;
    if (n < 1 || (n & (n - 1))) return set_error(AFFT_ERR_INVALID_LENGTH, "fast plans need a power of two >= 1");
    if (batch < 1) return set_error(AFFT_ERR_INVALID_LENGTH, "batch must be >= 1");
    if (!ctx || !out) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL handle");
    if (precision != AFFT_COMPLEX64 && precision != AFFT_COMPLEX128)
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "unknown precision");
    if (n > max_single_pass(precision))
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "no single-launch kernel for this length");
    afft_plan_desc d{};
    d.kind = AFFT_C2C;
    d.precision = precision;
    d.rank = 1;
    d.dims[0] = n;
    d.batch = batch;
    afft_plan p = nullptr;
    int rc = afft_plan_create_desc(ctx, &d, &p);
    if (rc) return rc;
    auto direct = [](const std::vector<Step>& v) {
        return v.size() == 1 && v[0].in_role == ROLE_SRC && v[0].out_role == ROLE_DST;
    };
    if (!direct(p->steps[0]) || !direct(p->steps[1])) {
        afft_plan_destroy(p);
        return set_error(AFFT_ERR_NOT_IMPLEMENTED, "no single-launch kernel for this length");
    }
    auto* fp = new afft_fast_plan_s();
    fp->plan = p;
    fp->stream = stream ? (hipStream_t)stream : p->stream;
    for (int di = 0; di < 2; ++di) fp->step[di] = p->steps[di][0];
    *out = fp;
    clear_error();
    return AFFT_OK;
}

int64_t afft_fast_plan_len(afft_fast_plan fp) { return fp ? fp->plan->n : 0; }

void afft_fast_forward(afft_fast_plan fp, void* dst, const void* src) { fast_launch(fp, 0, dst, src); }

void afft_fast_inverse(afft_fast_plan fp, void* dst, const void* src) { fast_launch(fp, 1, dst, src); }

int afft_fast_plan_destroy(afft_fast_plan fp) {
    delete fp;
    clear_error();
    return AFFT_OK;
}

}  // extern "C"
