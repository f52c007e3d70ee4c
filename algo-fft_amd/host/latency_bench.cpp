// Per-call host cost of a small device-resident transform through the C ABI:
// the checked entry point (afft_plan_exec_device: validation, plan lock,
// step loop) against a fast plan (afft_fast_forward: one pre-resolved
// launch, plan_fast.go:21-133).  Prints one JSON object.
//
//   latency_bench [n] [calls]
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "algofft_hip.h"

static double now_us() {
    using namespace std::chrono;
    return duration<double, std::micro>(steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char** argv) {
    const long long n = argc > 1 ? atoll(argv[1]) : 256;
    const int calls = argc > 2 ? atoi(argv[2]) : 20000;
    afft_context ctx = nullptr;
    if (afft_context_create(0, &ctx)) {
        printf("{\"error\": \"%s\"}\n", afft_last_error());
        return 1;
    }
    afft_buffer bx = nullptr, by = nullptr;
    afft_stream st = nullptr;
    afft_plan plan = nullptr;
    afft_fast_plan fp = nullptr;
    afft_plan_options opt{0, 1, 0};
    if (afft_buffer_create(ctx, n, AFFT_COMPLEX64, &bx) || afft_buffer_create(ctx, n, AFFT_COMPLEX64, &by) ||
        afft_stream_create(ctx, &st) || afft_plan_create(ctx, n, AFFT_COMPLEX64, &opt, &plan) ||
        afft_fast_plan_create(ctx, n, 1, AFFT_COMPLEX64, afft_stream_native(st), &fp)) {
        printf("{\"error\": \"%s\"}\n", afft_last_error());
        return 1;
    }
    void* x = afft_buffer_device_ptr(bx);
    void* y = afft_buffer_device_ptr(by);
    void* s = afft_stream_native(st);
    // enqueue cost: `calls` launches back to back, one synchronisation
    auto enqueue = [&](bool fast) {
        for (int i = 0; i < 100; ++i) fast ? afft_fast_forward(fp, y, x) : (void)afft_plan_exec_device(plan, AFFT_FORWARD, y, x, s);
        afft_stream_synchronize(st);
        const double t0 = now_us();
        for (int i = 0; i < calls; ++i) fast ? afft_fast_forward(fp, y, x) : (void)afft_plan_exec_device(plan, AFFT_FORWARD, y, x, s);
        afft_stream_synchronize(st);
        return (now_us() - t0) / calls;
    };
    // round trip: launch and wait, one transform at a time
    auto round_trip = [&](bool fast) {
        const int k = calls / 10;
        const double t0 = now_us();
        for (int i = 0; i < k; ++i) {
            fast ? afft_fast_forward(fp, y, x) : (void)afft_plan_exec_device(plan, AFFT_FORWARD, y, x, s);
            afft_stream_synchronize(st);
        }
        return (now_us() - t0) / k;
    };
    const double ec = enqueue(false), ef = enqueue(true), rc = round_trip(false), rf = round_trip(true);
    printf("{\"n\": %lld, \"calls\": %d, \"enqueue_checked_us\": %.3f, \"enqueue_fast_us\": %.3f, "
           "\"round_trip_checked_us\": %.3f, \"round_trip_fast_us\": %.3f}\n",
           n, calls, ec, ef, rc, rf);
    afft_fast_plan_destroy(fp);
    afft_plan_destroy(plan);
    afft_stream_destroy(st);
    afft_buffer_destroy(bx);
    afft_buffer_destroy(by);
    afft_context_destroy(ctx);
    return 0;
}
