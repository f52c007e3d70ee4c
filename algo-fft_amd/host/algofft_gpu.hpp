// algofft::gpu — C++ mirror of algo-fft's `gpu` package over the algofft_hip
// C ABI.  Header-only; link with libalgofft_hip.so.
//
// Names, argument meaning and error behaviour follow the Go sources:
//   Backend / Context / Buffer / Stream / PlanImpl     gpu/backend.go:7-51
//   RegisterBackend / CurrentBackendInfo               gpu/backend.go:53-81
//   Plan<T>: NewPlan, Len, Precision, Forward, Inverse,
//            ForwardInPlace, InverseInPlace, Close      gpu/plan.go:7-170
//   PlanOptions, DeviceInfo, BackendInfo, PrecisionKind gpu/types.go:10-42
//   sentinel errors                                     gpu/errors.go, errors.go
// Go's (value, error) returns become an Error return with out-parameters;
// Go slices become Slice<T> (pointer + length, nullptr = nil).
// Extensions mirroring the CPU Plan API (plan.go:504-571,
// plan_strided.go:16-99, plan_real_constructors.go:8-43, plan_2d.go:44,
// plan_3d.go:42, plan_nd.go:43): ForwardBatch/InverseBatch,
// ForwardStrided/InverseStrided, NewPlanReal32/64, NewPlan2D/3D/ND.
#pragma once
#include <complex>
#include <cstdint>
#include <memory>
#include <mutex>
#include <optional>
#include <string>
#include <type_traits>
#include <vector>

#include "algofft_hip.h"

namespace algofft {
namespace gpu {

// ---- errors (gpu/errors.go; errors.go for stride/spectrum) --------------------
struct Error {
    int code = AFFT_OK;
    std::string detail;
    explicit operator bool() const { return code != AFFT_OK; }   // true = an error
    bool Is(const Error& sentinel) const { return code == sentinel.code; }
    std::string String() const {
        std::string s = afft_status_string(code);
        return detail.empty() ? s : s + ": " + detail;
    }
};
inline const Error Nil{};
inline const Error ErrNoBackend{AFFT_ERR_NO_BACKEND, ""};
inline const Error ErrBackendUnavailable{AFFT_ERR_BACKEND_UNAVAILABLE, ""};
inline const Error ErrNotImplemented{AFFT_ERR_NOT_IMPLEMENTED, ""};
inline const Error ErrInvalidLength{AFFT_ERR_INVALID_LENGTH, ""};
inline const Error ErrNilSlice{AFFT_ERR_NIL_SLICE, ""};
inline const Error ErrLengthMismatch{AFFT_ERR_LENGTH_MISMATCH, ""};
inline const Error ErrInvalidStride{AFFT_ERR_INVALID_STRIDE, ""};
inline const Error ErrInvalidSpectrum{AFFT_ERR_INVALID_SPECTRUM, ""};

inline Error from_status(int rc) {
    if (rc == AFFT_OK) return Nil;
    return Error{rc, afft_last_error()};
}

// ---- types (gpu/types.go) ----------------------------------------------------
enum class PrecisionKind : uint8_t { Complex64 = AFFT_COMPLEX64, Complex128 = AFFT_COMPLEX128 };

struct DeviceInfo {
    std::string Name, Vendor, Driver;
    int64_t MemoryMB = 0;
    std::string ComputeCap;
};

struct BackendInfo {
    std::string Name, Version, Description;
};

struct PlanOptions {
    int DeviceIndex = 0;   // gpu/types.go:35
    int StreamCount = 0;   // :38
    bool InPlace = false;  // :41
    // extensions (plan_options.go:58-62)
    int64_t Batch = 0;
    int64_t Stride = 0;
};

template <typename T>
struct Slice {
    T* data = nullptr;
    int64_t len = 0;
    Slice() = default;
    Slice(T* d, int64_t n) : data(d), len(n) {}
    Slice(std::vector<std::remove_const_t<T>>& v) : data(v.data()), len((int64_t)v.size()) {}
    Slice(const std::vector<std::remove_const_t<T>>& v) : data(v.data()), len((int64_t)v.size()) {}
    bool nil() const { return data == nullptr; }
};

template <typename T>
struct precision_of;
template <>
struct precision_of<std::complex<float>> {
    static constexpr PrecisionKind value = PrecisionKind::Complex64;
};
template <>
struct precision_of<std::complex<double>> {
    static constexpr PrecisionKind value = PrecisionKind::Complex128;
};

// ---- interfaces (gpu/backend.go:7-51) ------------------------------------------
class Buffer {
   public:
    virtual ~Buffer() = default;
    virtual int64_t Len() const = 0;
    virtual PrecisionKind Precision() const = 0;
    virtual Error Upload(const void* src, int64_t len) = 0;
    virtual Error Download(void* dst, int64_t len) = 0;
    virtual Error Close() = 0;
};

class Stream {
   public:
    virtual ~Stream() = default;
    virtual Error Synchronize() = 0;
    virtual Error Close() = 0;
};

class PlanImpl {
   public:
    virtual ~PlanImpl() = default;
    virtual int64_t Len() const = 0;
    virtual PrecisionKind Precision() const = 0;
    // dst/src are host arrays of the plan's element type; lengths in elements
    virtual Error Forward(void* dst, int64_t dst_len, const void* src, int64_t src_len) = 0;
    virtual Error Inverse(void* dst, int64_t dst_len, const void* src, int64_t src_len) = 0;
    virtual Error Close() = 0;
};

class Context {
   public:
    virtual ~Context() = default;
    virtual DeviceInfo Device() const = 0;
    virtual Error NewBuffer(int64_t elemCount, PrecisionKind precision, std::unique_ptr<Buffer>* out) = 0;
    virtual Error NewStream(std::unique_ptr<Stream>* out) = 0;
    virtual Error NewFFTPlan(int64_t n, PrecisionKind precision, const PlanOptions& opts,
                             std::unique_ptr<PlanImpl>* out) = 0;
    // extended descriptor (batch / stride / real / N-D)
    virtual Error NewFFTPlanDesc(const afft_plan_desc& desc, std::unique_ptr<PlanImpl>* out) = 0;
    virtual Error Close() = 0;
};

class Backend {
   public:
    virtual ~Backend() = default;
    virtual BackendInfo Info() const = 0;
    virtual bool Available() const = 0;
    virtual Error Devices(std::vector<DeviceInfo>* out) const = 0;
    virtual Error NewContext(int deviceIndex, std::unique_ptr<Context>* out) = 0;
};

// ---- registry (gpu/backend.go:53-81) --------------------------------------------
namespace detail {
inline std::mutex& reg_mu() {
    static std::mutex m;
    return m;
}
inline std::shared_ptr<Backend>& reg() {
    static std::shared_ptr<Backend> b;
    return b;
}
}  // namespace detail

inline void RegisterBackend(std::shared_ptr<Backend> b) {
    std::lock_guard<std::mutex> l(detail::reg_mu());
    detail::reg() = std::move(b);
}
inline std::shared_ptr<Backend> getBackend() {
    std::lock_guard<std::mutex> l(detail::reg_mu());
    return detail::reg();
}
inline std::optional<BackendInfo> CurrentBackendInfo() {
    auto b = getBackend();
    if (!b) return std::nullopt;
    return b->Info();
}

// ---- the HIP backend ------------------------------------------------------------
class HIPBuffer : public Buffer {
   public:
    explicit HIPBuffer(afft_buffer h) : h_(h) {}
    ~HIPBuffer() override { Close(); }
    int64_t Len() const override { return afft_buffer_len(h_); }
    PrecisionKind Precision() const override { return (PrecisionKind)afft_buffer_precision(h_); }
    Error Upload(const void* src, int64_t len) override { return from_status(afft_buffer_upload(h_, src, len)); }
    Error Download(void* dst, int64_t len) override { return from_status(afft_buffer_download(h_, dst, len)); }
    Error Close() override {
        if (h_) afft_buffer_destroy(h_);
        h_ = nullptr;
        return Nil;
    }
    afft_buffer handle() const { return h_; }

   private:
    afft_buffer h_;
};

class HIPStream : public Stream {
   public:
    explicit HIPStream(afft_stream h) : h_(h) {}
    ~HIPStream() override { Close(); }
    Error Synchronize() override { return from_status(afft_stream_synchronize(h_)); }
    Error Close() override {
        if (h_) afft_stream_destroy(h_);
        h_ = nullptr;
        return Nil;
    }

   private:
    afft_stream h_;
};

class HIPPlanImpl : public PlanImpl {
   public:
    explicit HIPPlanImpl(afft_plan h) : h_(h) {}
    ~HIPPlanImpl() override { Close(); }
    int64_t Len() const override { return afft_plan_len(h_); }
    PrecisionKind Precision() const override { return (PrecisionKind)afft_plan_precision(h_); }
    Error Forward(void* dst, int64_t dl, const void* src, int64_t sl) override {
        return from_status(afft_plan_forward(h_, dst, dl, src, sl));
    }
    Error Inverse(void* dst, int64_t dl, const void* src, int64_t sl) override {
        return from_status(afft_plan_inverse(h_, dst, dl, src, sl));
    }
    Error Close() override {
        if (h_) afft_plan_destroy(h_);
        h_ = nullptr;
        return Nil;
    }
    std::string Algorithm() const {
        char buf[256];
        afft_plan_describe(h_, buf, sizeof(buf));
        return buf;
    }
    afft_plan handle() const { return h_; }

   private:
    afft_plan h_;
};

inline DeviceInfo to_info(const afft_device_info& d) {
    return DeviceInfo{d.name, d.vendor, d.driver, d.memory_mb, d.compute_cap};
}

class HIPContext : public Context {
   public:
    explicit HIPContext(afft_context h) : h_(h) {}
    ~HIPContext() override { Close(); }
    DeviceInfo Device() const override {
        afft_device_info d{};
        afft_context_device(h_, &d);
        return to_info(d);
    }
    Error NewBuffer(int64_t n, PrecisionKind p, std::unique_ptr<Buffer>* out) override {
        afft_buffer b = nullptr;
        Error e = from_status(afft_buffer_create(h_, n, (int)p, &b));
        if (!e) *out = std::make_unique<HIPBuffer>(b);
        return e;
    }
    Error NewStream(std::unique_ptr<Stream>* out) override {
        afft_stream s = nullptr;
        Error e = from_status(afft_stream_create(h_, &s));
        if (!e) *out = std::make_unique<HIPStream>(s);
        return e;
    }
    Error NewFFTPlan(int64_t n, PrecisionKind p, const PlanOptions& o, std::unique_ptr<PlanImpl>* out) override {
        afft_plan_desc d{};
        d.kind = AFFT_C2C;
        d.precision = (int)p;
        d.rank = 1;
        d.dims[0] = n;
        d.batch = o.Batch > 0 ? o.Batch : 1;
        d.istride = d.ostride = o.Stride > 0 ? o.Stride : 0;
        if (o.Stride > 1) d.idist = d.odist = 1;   // batch of interleaved columns
        if (o.InPlace) d.flags |= AFFT_FLAG_INPLACE;
        if (n < 1) return ErrInvalidLength;
        return NewFFTPlanDesc(d, out);
    }
    Error NewFFTPlanDesc(const afft_plan_desc& d, std::unique_ptr<PlanImpl>* out) override {
        afft_plan pl = nullptr;
        Error e = from_status(afft_plan_create_desc(h_, &d, &pl));
        if (!e) *out = std::make_unique<HIPPlanImpl>(pl);
        return e;
    }
    Error Close() override {
        if (h_) afft_context_destroy(h_);
        h_ = nullptr;
        return Nil;
    }

   private:
    afft_context h_;
};

class HIPBackend : public Backend {
   public:
    BackendInfo Info() const override {
        afft_backend_info i{};
        afft_backend_get_info(&i);
        return BackendInfo{i.name, i.version, i.description};
    }
    bool Available() const override { return afft_backend_available() != 0; }
    Error Devices(std::vector<DeviceInfo>* out) const override {
        int n = 0;
        Error e = from_status(afft_device_count(&n));
        if (e) return e;
        out->clear();
        for (int i = 0; i < n; ++i) {
            afft_device_info d{};
            if ((e = from_status(afft_device_get_info(i, &d)))) return e;
            out->push_back(to_info(d));
        }
        return Nil;
    }
    Error NewContext(int deviceIndex, std::unique_ptr<Context>* out) override {
        afft_context c = nullptr;
        Error e = from_status(afft_context_create(deviceIndex, &c));
        if (!e) *out = std::make_unique<HIPContext>(c);
        return e;
    }
};

// Opt-in registration, as the reference's build-tagged backends do
// (gpu/cuda_backend.go:29-32).
inline void RegisterHIPBackend() { RegisterBackend(std::make_shared<HIPBackend>()); }

// ---- Plan[T] (gpu/plan.go) -------------------------------------------------------
// Shared constructor body: gpu/plan.go:17-80 (checks in the same order, and
// the same unwinding of streams and context on failure).
inline Error open_plan(int64_t n, const PlanOptions& opts, std::unique_ptr<Context>* ctx,
                       std::vector<std::unique_ptr<Stream>>* streams) {
    if (n < 1) return ErrInvalidLength;
    auto backend = getBackend();
    if (!backend) return ErrNoBackend;
    if (!backend->Available()) return ErrBackendUnavailable;
    Error e = backend->NewContext(opts.DeviceIndex, ctx);
    if (e) return e;
    const int count = opts.StreamCount > 0 ? opts.StreamCount : 1;
    for (int i = 0; i < count; ++i) {
        std::unique_ptr<Stream> s;
        if ((e = (*ctx)->NewStream(&s))) {
            for (auto& t : *streams) t->Close();
            streams->clear();
            (*ctx)->Close();
            return e;
        }
        streams->push_back(std::move(s));
    }
    return Nil;
}

template <typename T>
class Plan {
   public:
    using value_type = T;

    // gpu.NewPlan[T](n, opts)
    static Error NewPlan(int64_t n, const PlanOptions& opts, std::unique_ptr<Plan>* out) {
        std::unique_ptr<Plan> p(new Plan());
        Error e = open_plan(n, opts, &p->ctx_, &p->streams_);
        if (e) return e;
        p->n_ = n;
        p->opts_ = opts;
        if ((e = p->ctx_->NewFFTPlan(n, precision_of<T>::value, opts, &p->impl_))) {
            p->closeStreamsAndCtx();
            return e;
        }
        *out = std::move(p);
        return Nil;
    }

    int64_t Len() const { return n_; }
    PrecisionKind Precision() const { return precision_of<T>::value; }

    // gpu/plan.go:100-114: nil -> ErrNilSlice; len < n -> ErrLengthMismatch
    Error Forward(Slice<T> dst, Slice<const T> src) { return run(dst, src, false); }
    Error Inverse(Slice<T> dst, Slice<const T> src) { return run(dst, src, true); }
    Error ForwardInPlace(Slice<T> data) { return Forward(data, Slice<const T>(data.data, data.len)); }
    Error InverseInPlace(Slice<T> data) { return Inverse(data, Slice<const T>(data.data, data.len)); }

    // Plan.ForwardBatch (plan.go:504-530): count >= 1, transform i at [i*n, (i+1)*n)
    Error ForwardBatch(Slice<T> dst, Slice<const T> src, int64_t count) { return batch(dst, src, count, false); }
    Error InverseBatch(Slice<T> dst, Slice<const T> src, int64_t count) { return batch(dst, src, count, true); }

    // Plan.ForwardStrided (plan_strided.go:16-135)
    Error ForwardStrided(Slice<T> dst, Slice<const T> src, int64_t stride) { return strided(dst, src, stride, false); }
    Error InverseStrided(Slice<T> dst, Slice<const T> src, int64_t stride) { return strided(dst, src, stride, true); }

    std::string Algorithm() const {
        auto* h = dynamic_cast<HIPPlanImpl*>(impl_.get());
        return h ? h->Algorithm() : "";
    }

    // gpu/plan.go:145-170: impl, then streams, then context; first error wins
    Error Close() {
        if (impl_) impl_->Close();
        impl_.reset();
        return closeStreamsAndCtx();
    }
    ~Plan() { Close(); }

   private:
    Plan() = default;
    Error closeStreamsAndCtx() {
        Error first;
        for (auto& s : streams_) {
            Error e = s->Close();
            if (e && !first) first = e;
        }
        streams_.clear();
        if (ctx_) {
            Error e = ctx_->Close();
            if (e && !first) first = e;
            ctx_.reset();
        }
        return first;
    }
    Error run(Slice<T> dst, Slice<const T> src, bool inv) {
        if (dst.nil() || src.nil()) return ErrNilSlice;
        if (dst.len < n_ || src.len < n_) return ErrLengthMismatch;
        if (!impl_) return ErrNotImplemented;
        return inv ? impl_->Inverse(dst.data, dst.len, src.data, src.len)
                   : impl_->Forward(dst.data, dst.len, src.data, src.len);
    }
    Error batch(Slice<T> dst, Slice<const T> src, int64_t count, bool inv) {
        if (dst.nil() || src.nil()) return ErrNilSlice;
        if (count < 1) return ErrInvalidLength;
        if (dst.len < count * n_ || src.len < count * n_) return ErrLengthMismatch;
        afft_plan_desc d = desc1(count, 0, 0);
        return adhoc(d, dst, src, inv);
    }
    Error strided(Slice<T> dst, Slice<const T> src, int64_t stride, bool inv) {
        if (dst.nil() || src.nil()) return ErrNilSlice;
        if (stride < 1) return ErrInvalidStride;
        if (stride == 1) return run(dst, src, inv);
        if (n_ - 1 > (INT64_MAX - 1) / stride) return ErrInvalidStride;
        const int64_t need = 1 + (n_ - 1) * stride;
        if (dst.len < need || src.len < need) return ErrLengthMismatch;
        afft_plan_desc d = desc1(1, stride, stride);
        return adhoc(d, dst, src, inv);
    }
    afft_plan_desc desc1(int64_t batch, int64_t is, int64_t os) const {
        afft_plan_desc d{};
        d.kind = AFFT_C2C;
        d.precision = (int)precision_of<T>::value;
        d.rank = 1;
        d.dims[0] = n_;
        d.batch = batch;
        d.istride = is;
        d.ostride = os;
        return d;
    }
    // batch and stride layouts get their own cached device plan
    Error adhoc(const afft_plan_desc& d, Slice<T> dst, Slice<const T> src, bool inv) {
        std::lock_guard<std::mutex> l(mu_);
        const std::string key = std::to_string(d.batch) + ":" + std::to_string(d.istride);
        if (key != cache_key_) {
            cache_.reset();
            Error e = ctx_->NewFFTPlanDesc(d, &cache_);
            if (e) return e;
            cache_key_ = key;
        }
        return inv ? cache_->Inverse(dst.data, dst.len, src.data, src.len)
                   : cache_->Forward(dst.data, dst.len, src.data, src.len);
    }

    int64_t n_ = 0;
    PlanOptions opts_;
    std::unique_ptr<Context> ctx_;
    std::vector<std::unique_ptr<Stream>> streams_;
    std::unique_ptr<PlanImpl> impl_;
    std::mutex mu_;
    std::unique_ptr<PlanImpl> cache_;
    std::string cache_key_;
};

// ---- descriptor plans: real and N-D ------------------------------------------------
// A plan built from a full descriptor.  Real plans (kind R2C) expose Forward
// (R2C, reals in, N/2+1 bins out) and Inverse (C2R), as PlanRealT does.
class DescPlan {
   public:
    static Error New(const afft_plan_desc& d, const PlanOptions& opts, std::unique_ptr<DescPlan>* out) {
        int64_t n = 1;
        for (int i = 0; i < d.rank; ++i) n *= d.dims[i];
        if (d.rank < 1 || n < 1) return ErrInvalidLength;
        if ((d.kind == AFFT_R2C || d.kind == AFFT_C2R) && (d.dims[d.rank - 1] < 2 || d.dims[d.rank - 1] % 2))
            return ErrInvalidLength;   // plan_real_generic.go:54-56
        std::unique_ptr<DescPlan> p(new DescPlan());
        Error e = open_plan(n, opts, &p->ctx_, &p->streams_);
        if (e) return e;
        if ((e = p->ctx_->NewFFTPlanDesc(d, &p->impl_))) return e;
        p->desc_ = d;
        *out = std::move(p);
        return Nil;
    }
    template <typename D, typename S>
    Error Forward(Slice<D> dst, Slice<const S> src) {
        if (dst.nil() || src.nil()) return ErrNilSlice;
        return impl_->Forward(dst.data, dst.len, src.data, src.len);
    }
    template <typename D, typename S>
    Error Inverse(Slice<D> dst, Slice<const S> src) {
        if (dst.nil() || src.nil()) return ErrNilSlice;
        return impl_->Inverse(dst.data, dst.len, src.data, src.len);
    }
    // PlanReal2D/3D.ForwardFull and InverseFull (plan_real_2d.go:231-276,
    // :370-389; plan_real_3d.go:241-285, :392-419): real plans only; the
    // full spectrum has as many bins as the input has points.
    template <typename D, typename S>
    Error ForwardFull(Slice<D> dst, Slice<const S> src) {
        if (dst.nil() || src.nil()) return ErrNilSlice;
        auto* h = dynamic_cast<HIPPlanImpl*>(impl_.get());
        if (!h) return ErrNotImplemented;
        return from_status(afft_plan_forward_full(h->handle(), dst.data, dst.len, src.data, src.len));
    }
    template <typename D, typename S>
    Error InverseFull(Slice<D> dst, Slice<const S> src) {
        if (dst.nil() || src.nil()) return ErrNilSlice;
        auto* h = dynamic_cast<HIPPlanImpl*>(impl_.get());
        if (!h) return ErrNotImplemented;
        return from_status(afft_plan_inverse_full(h->handle(), dst.data, dst.len, src.data, src.len));
    }
    Error Close() {
        impl_.reset();
        streams_.clear();
        ctx_.reset();
        return Nil;
    }

   private:
    DescPlan() = default;
    afft_plan_desc desc_{};
    std::unique_ptr<Context> ctx_;
    std::vector<std::unique_ptr<Stream>> streams_;
    std::unique_ptr<PlanImpl> impl_;
};

inline afft_plan_desc nd_desc(int kind, PrecisionKind p, const std::vector<int64_t>& dims, int64_t batch = 1) {
    afft_plan_desc d{};
    d.kind = kind;
    d.precision = (int)p;
    d.rank = (int)dims.size();
    for (size_t i = 0; i < dims.size() && i < 8; ++i) d.dims[i] = dims[i];
    d.batch = batch;
    return d;
}

// NewPlanReal32 / NewPlanReal64 (plan_real_constructors.go:8-43)
inline Error NewPlanReal32(int64_t n, std::unique_ptr<DescPlan>* out, const PlanOptions& o = {}) {
    return DescPlan::New(nd_desc(AFFT_R2C, PrecisionKind::Complex64, {n}, o.Batch > 0 ? o.Batch : 1), o, out);
}
inline Error NewPlanReal64(int64_t n, std::unique_ptr<DescPlan>* out, const PlanOptions& o = {}) {
    return DescPlan::New(nd_desc(AFFT_R2C, PrecisionKind::Complex128, {n}, o.Batch > 0 ? o.Batch : 1), o, out);
}
// NewPlan2D / NewPlan3D / NewPlanND (plan_2d.go:44, plan_3d.go:42, plan_nd.go:43)
template <typename T>
inline Error NewPlan2D(int64_t rows, int64_t cols, std::unique_ptr<DescPlan>* out, const PlanOptions& o = {}) {
    return DescPlan::New(nd_desc(AFFT_C2C, precision_of<T>::value, {rows, cols}), o, out);
}
template <typename T>
inline Error NewPlan3D(int64_t d, int64_t h, int64_t w, std::unique_ptr<DescPlan>* out, const PlanOptions& o = {}) {
    return DescPlan::New(nd_desc(AFFT_C2C, precision_of<T>::value, {d, h, w}), o, out);
}
template <typename T>
inline Error NewPlanND(const std::vector<int64_t>& dims, std::unique_ptr<DescPlan>* out, const PlanOptions& o = {}) {
    if (dims.empty() || dims.size() > 8) return ErrInvalidLength;
    return DescPlan::New(nd_desc(AFFT_C2C, precision_of<T>::value, dims), o, out);
}
// NewPlanReal2D / NewPlanReal3D (plan_real_2d.go:43, plan_real_3d.go:42)
inline Error NewPlanReal2D(int64_t rows, int64_t cols, std::unique_ptr<DescPlan>* out, const PlanOptions& o = {}) {
    return DescPlan::New(nd_desc(AFFT_R2C, PrecisionKind::Complex64, {rows, cols}), o, out);
}
inline Error NewPlanReal3D(int64_t d, int64_t h, int64_t w, std::unique_ptr<DescPlan>* out,
                           const PlanOptions& o = {}) {
    return DescPlan::New(nd_desc(AFFT_R2C, PrecisionKind::Complex64, {d, h, w}), o, out);
}

// ---- convolution and correlation on the device --------------------------------
// Convolve / Convolve128 (convolve.go:9-115), ConvolveReal (convolve_real.go:9-80),
// CrossCorrelate / AutoCorrelate (correlate.go:6-79): the reference's checks in
// its order (nil, empty, dst length), then one synchronised device call on a
// context of `device`.
namespace detail {
inline Error convolve_on(int device, PrecisionKind p, int mode, void* dst, int64_t dl, const void* a, int64_t na,
                         const void* b, int64_t nb) {
    if (!dst || !a || !b) return ErrNilSlice;
    if (na < 1 || nb < 1) return ErrInvalidLength;
    if (dl != na + nb - 1) return ErrLengthMismatch;
    afft_context ctx = nullptr;
    Error e = from_status(afft_context_create(device, &ctx));
    if (e) return e;
    e = from_status(afft_convolve_host(ctx, (int)p, mode, dst, dl, a, na, b, nb));
    afft_context_destroy(ctx);
    return e;
}
}  // namespace detail

template <typename T>
inline Error Convolve(Slice<T> dst, Slice<const T> a, Slice<const T> b, int device = 0) {
    return detail::convolve_on(device, precision_of<T>::value, AFFT_CONV_LINEAR, dst.data, dst.len, a.data, a.len,
                               b.data, b.len);
}
template <typename T>
inline Error CrossCorrelate(Slice<T> dst, Slice<const T> a, Slice<const T> b, int device = 0) {
    return detail::convolve_on(device, precision_of<T>::value, AFFT_CONV_CORRELATE, dst.data, dst.len, a.data, a.len,
                               b.data, b.len);
}
template <typename T>
inline Error AutoCorrelate(Slice<T> dst, Slice<const T> a, int device = 0) {
    return CrossCorrelate<T>(dst, a, a, device);
}
inline Error ConvolveReal(Slice<float> dst, Slice<const float> a, Slice<const float> b, int device = 0) {
    return detail::convolve_on(device, PrecisionKind::Complex64, AFFT_CONV_REAL, dst.data, dst.len, a.data, a.len,
                               b.data, b.len);
}
inline Error ConvolveReal64(Slice<double> dst, Slice<const double> a, Slice<const double> b, int device = 0) {
    return detail::convolve_on(device, PrecisionKind::Complex128, AFFT_CONV_REAL, dst.data, dst.len, a.data, a.len,
                               b.data, b.len);
}

}  // namespace gpu
}  // namespace algofft
