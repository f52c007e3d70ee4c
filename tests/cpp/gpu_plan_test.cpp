// Tests of the C++ mirror of algo-fft's gpu package (algofft_gpu.hpp), in the
// style of gpu/mock_backend_test.go and the root plan tests.
//   gpu_plan_test cpu   error paths that need no device (gpu/plan.go:17-29)
//   gpu_plan_test gpu   transforms on device 0, checked against a direct DFT
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <vector>

#include "algofft_gpu.hpp"

using namespace algofft::gpu;
using c64 = std::complex<float>;
using c128 = std::complex<double>;

static int failures = 0;
#define CHECK(cond, ...)                                      \
    do {                                                      \
        if (!(cond)) {                                        \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__);  \
            std::printf(__VA_ARGS__);                         \
            std::printf("\n");                                \
            ++failures;                                       \
        }                                                     \
    } while (0)

template <typename T>
static std::vector<c128> dft(const std::vector<T>& x, int sign) {
    const size_t n = x.size();
    std::vector<c128> y(n);
    for (size_t k = 0; k < n; ++k) {
        long double sr = 0, si = 0;
        for (size_t j = 0; j < n; ++j) {
            const long double a = sign * 2.0L * 3.141592653589793238462643383279L * (long double)((k * j) % n) / n;
            const long double xr = std::real(x[j]), xi = std::imag(x[j]);
            sr += xr * cosl(a) - xi * sinl(a);
            si += xr * sinl(a) + xi * cosl(a);
        }
        y[k] = c128((double)sr, (double)si);
    }
    return y;
}

template <typename A, typename B>
static double rel_l2(const std::vector<A>& a, const std::vector<B>& b) {
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        const c128 d = c128(a[i]) - c128(b[i]);
        num += std::norm(d);
        den += std::norm(c128(b[i]));
    }
    return std::sqrt(num / (den > 0 ? den : 1));
}

static void cpu_tests() {
    // no backend registered -> ErrNoBackend (gpu/plan.go:22-25)
    RegisterBackend(nullptr);
    std::unique_ptr<Plan<c64>> p;
    Error e = Plan<c64>::NewPlan(64, {}, &p);
    CHECK(e.Is(ErrNoBackend), "want ErrNoBackend, got %s", e.String().c_str());
    CHECK(!CurrentBackendInfo().has_value(), "no backend info expected");
    // invalid length is checked first (gpu/plan.go:18-20)
    RegisterHIPBackend();
    e = Plan<c64>::NewPlan(0, {}, &p);
    CHECK(e.Is(ErrInvalidLength), "want ErrInvalidLength, got %s", e.String().c_str());
    CHECK(CurrentBackendInfo()->Name == "hip", "backend name");
    // without a device: ErrBackendUnavailable (gpu/plan.go:27-29)
    auto b = getBackend();
    if (!b->Available()) {
        e = Plan<c64>::NewPlan(64, {}, &p);
        CHECK(e.Is(ErrBackendUnavailable), "want ErrBackendUnavailable, got %s", e.String().c_str());
    }
    // real plans reject odd lengths (plan_real_generic.go:54-56)
    std::unique_ptr<DescPlan> rp;
    e = NewPlanReal32(7, &rp);
    CHECK(e.Is(ErrInvalidLength), "odd real length: %s", e.String().c_str());
}

static void gpu_tests() {
    RegisterHIPBackend();
    std::vector<DeviceInfo> devs;
    CHECK(!getBackend()->Devices(&devs) && !devs.empty(), "devices");
    // impulse -> ones; round trip (mock_backend_test.go:5-23 with values)
    {
        std::unique_ptr<Plan<c64>> p;
        Error e = Plan<c64>::NewPlan(1024, {}, &p);
        CHECK(!e, "NewPlan: %s", e.String().c_str());
        std::vector<c64> x(1024), X(1024), y(1024);
        x[0] = 1;
        CHECK(!p->Forward(X, x), "Forward");
        double m = 0;
        for (auto& v : X) m = std::fmax(m, std::abs(v - c64(1)));
        CHECK(m < 1e-6, "impulse spectrum %g", m);
        CHECK(!p->Inverse(y, X), "Inverse");
        CHECK(rel_l2(y, x) < 1e-6, "round trip");
        // nil and short slices
        CHECK(p->Forward(Slice<c64>(), x).Is(ErrNilSlice), "nil dst");
        std::vector<c64> shortv(10);
        CHECK(p->Forward(shortv, x).Is(ErrLengthMismatch), "short dst");
        CHECK(p->Algorithm().find("hip_") == 0, "algorithm string");
        CHECK(!p->Close(), "Close");
    }
    // complex128 mixed radix and Bluestein against a direct DFT
    for (int64_t n : {360, 509}) {
        std::unique_ptr<Plan<c128>> p;
        CHECK(!Plan<c128>::NewPlan(n, {}, &p), "NewPlan %lld", (long long)n);
        std::vector<c128> x(n), X(n);
        for (int64_t i = 0; i < n; ++i) x[i] = c128(std::sin(0.3 * i), std::cos(1.7 * i));
        CHECK(!p->Forward(X, x), "Forward");
        CHECK(rel_l2(X, dft(x, -1)) < 1e-13, "c128 n=%lld err %g", (long long)n, rel_l2(X, dft(x, -1)));
    }
    // ForwardBatch == loop of Forward (plan_batch_test.go:13-146)
    {
        const int64_t n = 256, cnt = 7;
        std::unique_ptr<Plan<c64>> p;
        CHECK(!Plan<c64>::NewPlan(n, {}, &p), "NewPlan");
        std::vector<c64> x(n * cnt), X(n * cnt), one(n);
        for (size_t i = 0; i < x.size(); ++i) x[i] = c64(std::sin(0.1f * i), std::cos(0.37f * i));
        CHECK(!p->ForwardBatch(X, x, cnt), "ForwardBatch");
        CHECK(p->ForwardBatch(X, x, 0).Is(ErrInvalidLength), "count 0");
        for (int64_t b = 0; b < cnt; ++b) {
            std::vector<c64> xb(x.begin() + b * n, x.begin() + (b + 1) * n);
            CHECK(!p->Forward(one, xb), "Forward");
            CHECK(std::memcmp(one.data(), X.data() + b * n, n * sizeof(c64)) == 0, "batch %lld differs", (long long)b);
        }
        // strided column of a 256 x 3 matrix (plan_strided.go:35-99)
        std::vector<c64> m(n * 3), out(n * 3);
        for (size_t i = 0; i < m.size(); ++i) m[i] = c64((float)(i % 5), (float)(i % 3));
        CHECK(!p->ForwardStrided(out, m, 3), "ForwardStrided");
        CHECK(p->ForwardStrided(out, m, 0).Is(ErrInvalidStride), "stride 0");
        std::vector<c64> col(n), colX(n);
        for (int64_t i = 0; i < n; ++i) col[i] = m[i * 3];
        p->Forward(colX, col);
        std::vector<c64> got(n);
        for (int64_t i = 0; i < n; ++i) got[i] = out[i * 3];
        CHECK(rel_l2(got, colX) < 1e-6, "strided vs gathered");
    }
    // NewPlanReal32 forward / inverse
    {
        const int64_t n = 1000;
        std::unique_ptr<DescPlan> rp;
        CHECK(!NewPlanReal32(n, &rp), "NewPlanReal32");
        std::vector<float> x(n), y(n);
        for (int64_t i = 0; i < n; ++i) x[i] = std::sin(0.05f * i) + 0.25f * (i % 7);
        std::vector<c64> X(n / 2 + 1);
        CHECK(!rp->Forward(Slice<c64>(X), Slice<const float>(x)), "R2C");
        std::vector<c128> xc(x.begin(), x.end());
        std::vector<c128> full = dft(xc, -1);
        full.resize(n / 2 + 1);
        CHECK(rel_l2(X, full) < 2e-6 * 10, "r2c err %g", rel_l2(X, full));
        CHECK(!rp->Inverse(Slice<float>(y), Slice<const c64>(X)), "C2R");
        std::vector<c64> yc(y.begin(), y.end()), xcc(x.begin(), x.end());
        CHECK(rel_l2(yc, xcc) < 4e-5, "c2r round trip");
    }
    // NewPlan2D against two passes of direct DFTs
    {
        const int64_t r = 12, c = 20;
        std::unique_ptr<DescPlan> p2;
        CHECK(!NewPlan2D<c128>(r, c, &p2), "NewPlan2D");
        std::vector<c128> x(r * c), X(r * c);
        for (size_t i = 0; i < x.size(); ++i) x[i] = c128(std::cos(0.7 * i), std::sin(0.2 * i));
        CHECK(!p2->Forward(Slice<c128>(X), Slice<const c128>(x)), "2-D forward");
        std::vector<c128> t(r * c), ref(r * c);
        for (int64_t i = 0; i < r; ++i) {
            std::vector<c128> row(x.begin() + i * c, x.begin() + (i + 1) * c);
            auto R = dft(row, -1);
            for (int64_t j = 0; j < c; ++j) t[i * c + j] = R[j];
        }
        for (int64_t j = 0; j < c; ++j) {
            std::vector<c128> col(r);
            for (int64_t i = 0; i < r; ++i) col[i] = t[i * c + j];
            auto C = dft(col, -1);
            for (int64_t i = 0; i < r; ++i) ref[i * c + j] = C[i];
        }
        CHECK(rel_l2(X, ref) < 1e-13, "2-D err %g", rel_l2(X, ref));
    }
    // PlanReal3D.ForwardFull / InverseFull as plan_real_3d_test.go:133-228
    // checks them (compact bins match Forward; InverseFull inverts), plus the
    // redundant bins against direct DFTs over all three axes
    {
        const int64_t d = 3, h = 4, w = 6, n = d * h * w, hw = w / 2 + 1;
        std::unique_ptr<DescPlan> rp;
        CHECK(!NewPlanReal3D(d, h, w, &rp), "NewPlanReal3D");
        std::vector<float> x(n), y(n);
        for (int64_t i = 0; i < n; ++i) x[i] = std::cos(0.37f * i) - 0.1f * (i % 5);
        std::vector<c64> X(d * h * hw), F(n);
        CHECK(!rp->Forward(Slice<c64>(X), Slice<const float>(x)), "3-D R2C");
        CHECK(!rp->ForwardFull(Slice<c64>(F), Slice<const float>(x)), "ForwardFull");
        double dev = 0;
        for (int64_t r = 0; r < d * h; ++r)
            for (int64_t c = 0; c < hw; ++c) dev = std::max(dev, (double)std::abs(F[r * w + c] - X[r * hw + c]));
        CHECK(dev < 1e-5, "ForwardFull compact bins differ by %g", dev);
        std::vector<c128> ref(n);
        for (int64_t k0 = 0; k0 < d; ++k0)
            for (int64_t k1 = 0; k1 < h; ++k1)
                for (int64_t k2 = 0; k2 < w; ++k2) {
                    c128 acc = 0;
                    for (int64_t i0 = 0; i0 < d; ++i0)
                        for (int64_t i1 = 0; i1 < h; ++i1)
                            for (int64_t i2 = 0; i2 < w; ++i2) {
                                const double ph = -2 * M_PI * ((double)(k0 * i0 % d) / d + (double)(k1 * i1 % h) / h +
                                                               (double)(k2 * i2 % w) / w);
                                acc += (double)x[(i0 * h + i1) * w + i2] * c128(std::cos(ph), std::sin(ph));
                            }
                    ref[(k0 * h + k1) * w + k2] = acc;
                }
        CHECK(rel_l2(F, ref) < 2e-6 * 7, "ForwardFull err %g", rel_l2(F, ref));
        CHECK(!rp->InverseFull(Slice<float>(y), Slice<const c64>(F)), "InverseFull");
        std::vector<c64> yc(y.begin(), y.end()), xc(x.begin(), x.end());
        CHECK(rel_l2(yc, xc) < 4e-5, "InverseFull round trip %g", rel_l2(yc, xc));
        std::vector<c64> short_out(n - 1);
        CHECK(rp->ForwardFull(Slice<c64>(short_out), Slice<const float>(x)).Is(ErrLengthMismatch),
              "ForwardFull length check");
    }
}

// Convolve / ConvolveReal / CrossCorrelate / AutoCorrelate through the mirror
// (convolve_test.go:14 and convolve_real_test.go:10-29 known answers; checks
// before any device call in the cpu mode).
static void conv_tests(bool gpu) {
    std::vector<c64> a{c64(1, 0), c64(2, 0), c64(3, 0)}, b{c64(4, 0), c64(5, 0)}, d(4), shortd(3);
    CHECK(Convolve<c64>(Slice<c64>(), a, b).Is(ErrNilSlice), "nil dst");
    CHECK(Convolve<c64>(d, Slice<const c64>(a.data(), 0), b).Is(ErrInvalidLength), "empty a");
    CHECK(Convolve<c64>(shortd, a, b).Is(ErrLengthMismatch), "short dst");
    std::vector<float> ra{1, 2, 3}, rb{4, 5}, rd(4);
    CHECK(ConvolveReal(Slice<float>(), ra, rb).Is(ErrNilSlice), "real nil dst");
    if (!gpu) return;
    CHECK(!Convolve<c64>(d, a, b), "Convolve");
    const float want[4] = {4, 13, 22, 15};
    for (int i = 0; i < 4; ++i) CHECK(std::abs(d[i] - c64(want[i], 0)) < 1e-4f, "conv[%d]", i);
    CHECK(!ConvolveReal(rd, ra, rb), "ConvolveReal");
    for (int i = 0; i < 4; ++i) CHECK(std::fabs(rd[i] - want[i]) < 1e-4f, "real conv[%d]", i);
    // AutoCorrelate of [1, i]: lags -1, 0, 1 -> [-i, 2, i]
    std::vector<c128> x{c128(1, 0), c128(0, 1)}, ac(3);
    CHECK(!AutoCorrelate<c128>(ac, x), "AutoCorrelate");
    CHECK(std::abs(ac[0] - c128(0, -1)) < 1e-12 && std::abs(ac[1] - c128(2, 0)) < 1e-12 &&
              std::abs(ac[2] - c128(0, 1)) < 1e-12,
          "autocorrelation of [1, i]");
}

int main(int argc, char** argv) {
    const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
    cpu_tests();
    conv_tests(gpu);
    if (gpu) gpu_tests();
    std::printf("%s tests: %s (%d failures)\n", gpu ? "gpu" : "cpu", failures ? "FAIL" : "ok", failures);
    return failures ? 1 : 0;
}
