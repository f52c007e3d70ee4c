// Cost model: a 32 -> 64 -> 64 -> 1 MLP (ReLU hidden layers) predicting the
// standardised log runtime of a (shape, candidate) pair, trained with Adam on
// the GPU.  Forward, backward and the optimiser step are the kernels below.
//
// Parameter layout (floats): W1[32][64], b1[64], W2[64][64], b2[64], W3[64], b3.
// Loss: mean squared error over the batch.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "afft_internal.hpp"

namespace afft {

constexpr int MI = NFEAT, MH = 64;
constexpr int OFF_W1 = 0, OFF_B1 = MI * MH, OFF_W2 = OFF_B1 + MH, OFF_B2 = OFF_W2 + MH * MH, OFF_W3 = OFF_B2 + MH,
              OFF_B3 = OFF_W3 + MH, NPARAM = OFF_B3 + 1;

// One block (64 threads) per sample: hidden activations and the output.
__global__ void __launch_bounds__(64) k_mlp_fwd(const float* __restrict__ P, const float* __restrict__ X, int n,
                                                float* __restrict__ H1, float* __restrict__ H2, float* __restrict__ Y) {
    __shared__ float x[MI], h1[MH], red[MH];
    const int s = blockIdx.x, j = threadIdx.x;
    if (s >= n) return;
    if (j < MI) x[j] = X[(long long)s * MI + j];
    __syncthreads();
    float a = P[OFF_B1 + j];
    for (int i = 0; i < MI; ++i) a += x[i] * P[OFF_W1 + i * MH + j];
    a = a > 0.f ? a : 0.f;
    h1[j] = a;
    H1[(long long)s * MH + j] = a;
    __syncthreads();
    float b = P[OFF_B2 + j];
    for (int i = 0; i < MH; ++i) b += h1[i] * P[OFF_W2 + i * MH + j];
    b = b > 0.f ? b : 0.f;
    H2[(long long)s * MH + j] = b;
    red[j] = b * P[OFF_W3 + j];
    __syncthreads();
    for (int w = MH / 2; w > 0; w >>= 1) {
        if (j < w) red[j] += red[j + w];
        __syncthreads();
    }
    if (j == 0) Y[s] = red[0] + P[OFF_B3];
}

// Per-sample deltas: dY = 2(Y - T)/n, dH2 = dY*W3*[H2>0], dH1 = (W2 dH2)*[H1>0].
__global__ void __launch_bounds__(64) k_mlp_delta(const float* __restrict__ P, const float* __restrict__ T, int n,
                                                  const float* __restrict__ H1, const float* __restrict__ H2,
                                                  const float* __restrict__ Y, float* __restrict__ dY,
                                                  float* __restrict__ dH1, float* __restrict__ dH2) {
    __shared__ float d2[MH];
    const int s = blockIdx.x, j = threadIdx.x;
    if (s >= n) return;
    const float dy = 2.f * (Y[s] - T[s]) / (float)n;
    if (j == 0) dY[s] = dy;
    const float g2 = H2[(long long)s * MH + j] > 0.f ? dy * P[OFF_W3 + j] : 0.f;
    d2[j] = g2;
    dH2[(long long)s * MH + j] = g2;
    __syncthreads();
    float g1 = 0.f;
    for (int k = 0; k < MH; ++k) g1 += P[OFF_W2 + j * MH + k] * d2[k];
    dH1[(long long)s * MH + j] = H1[(long long)s * MH + j] > 0.f ? g1 : 0.f;
}

// One thread per parameter: its gradient summed over the batch.
__global__ void __launch_bounds__(256) k_mlp_wgrad(const float* __restrict__ X, int n, const float* __restrict__ H1,
                                                   const float* __restrict__ H2, const float* __restrict__ dY,
                                                   const float* __restrict__ dH1, const float* __restrict__ dH2,
                                                   float* __restrict__ G) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= NPARAM) return;
    float g = 0.f;
    if (p < OFF_B1) {
        const int i = p / MH, j = p % MH;
        for (int s = 0; s < n; ++s) g += X[(long long)s * MI + i] * dH1[(long long)s * MH + j];
    } else if (p < OFF_W2) {
        const int j = p - OFF_B1;
        for (int s = 0; s < n; ++s) g += dH1[(long long)s * MH + j];
    } else if (p < OFF_B2) {
        const int q = p - OFF_W2, i = q / MH, j = q % MH;
        for (int s = 0; s < n; ++s) g += H1[(long long)s * MH + i] * dH2[(long long)s * MH + j];
    } else if (p < OFF_W3) {
        const int j = p - OFF_B2;
        for (int s = 0; s < n; ++s) g += dH2[(long long)s * MH + j];
    } else if (p < OFF_B3) {
        const int j = p - OFF_W3;
        for (int s = 0; s < n; ++s) g += H2[(long long)s * MH + j] * dY[s];
    } else {
        for (int s = 0; s < n; ++s) g += dY[s];
    }
    G[p] = g;
}

__global__ void __launch_bounds__(256) k_adam(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M,
                                              float* __restrict__ V, float lr, float b1, float b2, float eps,
                                              float c1, float c2) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= NPARAM) return;
    const float g = G[p];
    const float m = b1 * M[p] + (1.f - b1) * g;
    const float v = b2 * V[p] + (1.f - b2) * g * g;
    M[p] = m;
    V[p] = v;
    P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

struct Mlp {
    int device = 0;
    float* P = nullptr;
    float* Mm = nullptr;
    float* Vv = nullptr;
    float* G = nullptr;
    int step = 0;
    float mean[NFEAT], stdv[NFEAT];
    float ymean = 0.f, ystd = 1.f;
    // work buffers sized for `cap` samples
    int64_t cap = 0;
    float *X = nullptr, *T = nullptr, *H1 = nullptr, *H2 = nullptr, *Y = nullptr, *dY = nullptr, *dH1 = nullptr,
          *dH2 = nullptr;
    ~Mlp() {
        (void)hipSetDevice(device);
        for (float* q : {P, Mm, Vv, G, X, T, H1, H2, Y, dY, dH1, dH2})
            if (q) (void)hipFree(q);
    }
    int ensure(int64_t n) {
        if (n <= cap) return AFFT_OK;
        for (float** q : {&X, &T, &H1, &H2, &Y, &dY, &dH1, &dH2}) {
            if (*q) (void)hipFree(*q);
            *q = nullptr;
        }
        cap = 0;
        AFFT_HIP(hipMalloc(&X, n * MI * 4));
        AFFT_HIP(hipMalloc(&T, n * 4));
        AFFT_HIP(hipMalloc(&H1, n * MH * 4));
        AFFT_HIP(hipMalloc(&H2, n * MH * 4));
        AFFT_HIP(hipMalloc(&Y, n * 4));
        AFFT_HIP(hipMalloc(&dY, n * 4));
        AFFT_HIP(hipMalloc(&dH1, n * MH * 4));
        AFFT_HIP(hipMalloc(&dH2, n * MH * 4));
        cap = n;
        return AFFT_OK;
    }
    void normalise(const float* x, int64_t n, std::vector<float>& out) const {
        out.resize(n * MI);
        for (int64_t s = 0; s < n; ++s)
            for (int i = 0; i < MI; ++i) out[s * MI + i] = (x[s * MI + i] - mean[i]) / stdv[i];
    }
    int forward(int64_t n) {
        void* a[] = {&P, &X, &n, &H1, &H2, &Y};
        int nn = (int)n;
        a[2] = &nn;
        AFFT_HIP(hipLaunchKernel((const void*)&k_mlp_fwd, dim3((unsigned)n), dim3(64), a, 0, nullptr));
        return AFFT_OK;
    }
    int backward(int64_t n) {
        int nn = (int)n;
        void* a[] = {&P, &T, &nn, &H1, &H2, &Y, &dY, &dH1, &dH2};
        AFFT_HIP(hipLaunchKernel((const void*)&k_mlp_delta, dim3((unsigned)n), dim3(64), a, 0, nullptr));
        void* b[] = {&X, &nn, &H1, &H2, &dY, &dH1, &dH2, &G};
        AFFT_HIP(hipLaunchKernel((const void*)&k_mlp_wgrad, dim3((NPARAM + 255) / 256), dim3(256), b, 0, nullptr));
        return AFFT_OK;
    }
};

int mlp_predict_host(Mlp* m, const float* x, int64_t n, float* y) {
    if (!m || n < 1) return AFFT_ERR_INVALID_ARGUMENT;
    AFFT_HIP(hipSetDevice(m->device));
    int rc = m->ensure(n);
    if (rc) return rc;
    std::vector<float> xn;
    m->normalise(x, n, xn);
    AFFT_HIP(hipMemcpy(m->X, xn.data(), xn.size() * 4, hipMemcpyHostToDevice));
    if ((rc = m->forward(n))) return rc;
    AFFT_HIP(hipMemcpy(y, m->Y, n * 4, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; ++i) y[i] = y[i] * m->ystd + m->ymean;
    return AFFT_OK;
}

}  // namespace afft

using namespace afft;

struct afft_mlp_s : Mlp {};

static uint64_t sm64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

extern "C" {

int afft_mlp_param_count(void) { return NPARAM; }

// He-uniform initial weights from a seeded stream, zero biases.
int afft_mlp_create(afft_context ctx, uint64_t seed, afft_mlp* out) {
    if (!ctx || !out) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    AFFT_HIP(hipSetDevice(ctx->device));
    auto* m = new afft_mlp_s();
    m->device = ctx->device;
    for (int i = 0; i < NFEAT; ++i) {
        m->mean[i] = 0.f;
        m->stdv[i] = 1.f;
    }
    std::vector<float> p(NPARAM, 0.f);
    uint64_t c = seed * 0xD1342543DE82EF95ULL;
    auto u = [&] { return (double)(sm64(c++) >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0; };
    for (int i = 0; i < MI * MH; ++i) p[OFF_W1 + i] = (float)(u() * std::sqrt(6.0 / MI));
    for (int i = 0; i < MH * MH; ++i) p[OFF_W2 + i] = (float)(u() * std::sqrt(6.0 / MH));
    for (int i = 0; i < MH; ++i) p[OFF_W3 + i] = (float)(u() * std::sqrt(6.0 / MH));
    hipError_t e = hipSuccess;
    for (float** q : {&m->P, &m->Mm, &m->Vv, &m->G})
        if (e == hipSuccess) e = hipMalloc(q, NPARAM * 4);
    if (e == hipSuccess) e = hipMemcpy(m->P, p.data(), NPARAM * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(m->Mm, 0, NPARAM * 4);
    if (e == hipSuccess) e = hipMemset(m->Vv, 0, NPARAM * 4);
    if (e != hipSuccess) {
        delete m;
        return set_error(AFFT_ERR_DEVICE, std::string("mlp alloc: ") + hipGetErrorString(e));
    }
    *out = m;
    clear_error();
    return AFFT_OK;
}

int afft_mlp_set_normalization(afft_mlp m, const float* mean, const float* stdv, float ymean, float ystd) {
    if (!m || !mean || !stdv) return set_error(AFFT_ERR_NIL_SLICE, "NULL argument");
    for (int i = 0; i < NFEAT; ++i) {
        m->mean[i] = mean[i];
        m->stdv[i] = stdv[i] > 1e-12f ? stdv[i] : 1.f;
    }
    m->ymean = ymean;
    m->ystd = ystd > 1e-12f ? ystd : 1.f;
    clear_error();
    return AFFT_OK;
}

int afft_mlp_get_params(afft_mlp m, float* out) {
    if (!m || !out) return set_error(AFFT_ERR_NIL_SLICE, "NULL argument");
    AFFT_HIP(hipSetDevice(m->device));
    AFFT_HIP(hipMemcpy(out, m->P, NPARAM * 4, hipMemcpyDeviceToHost));
    clear_error();
    return AFFT_OK;
}

int afft_mlp_set_params(afft_mlp m, const float* in) {
    if (!m || !in) return set_error(AFFT_ERR_NIL_SLICE, "NULL argument");
    AFFT_HIP(hipSetDevice(m->device));
    AFFT_HIP(hipMemcpy(m->P, in, NPARAM * 4, hipMemcpyHostToDevice));
    clear_error();
    return AFFT_OK;
}

// Loss and parameter gradient of a batch with raw targets (standardised
// inside like the features), for the gradient check.
int afft_mlp_loss_grad(afft_mlp m, const float* x, const float* t, int64_t n, float* loss, float* grad) {
    if (!m || !x || !t || n < 1) return set_error(AFFT_ERR_INVALID_ARGUMENT, "bad argument");
    AFFT_HIP(hipSetDevice(m->device));
    int rc = m->ensure(n);
    if (rc) return rc;
    std::vector<float> xn, tn(n);
    m->normalise(x, n, xn);
    for (int64_t i = 0; i < n; ++i) tn[i] = (t[i] - m->ymean) / m->ystd;
    AFFT_HIP(hipMemcpy(m->X, xn.data(), xn.size() * 4, hipMemcpyHostToDevice));
    AFFT_HIP(hipMemcpy(m->T, tn.data(), n * 4, hipMemcpyHostToDevice));
    if ((rc = m->forward(n)) || (rc = m->backward(n))) return rc;
    std::vector<float> y(n);
    AFFT_HIP(hipMemcpy(y.data(), m->Y, n * 4, hipMemcpyDeviceToHost));
    if (grad) AFFT_HIP(hipMemcpy(grad, m->G, NPARAM * 4, hipMemcpyDeviceToHost));
    if (loss) {
        double l = 0;
        for (int64_t i = 0; i < n; ++i) l += (double)(y[i] - tn[i]) * (y[i] - tn[i]);
        *loss = (float)(l / n);
    }
    clear_error();
    return AFFT_OK;
}

// Mini-batch Adam over `epochs` passes of a seeded shuffle.  Returns the
// final full-set loss (standardised units) in *final_loss.
int afft_mlp_train(afft_mlp m, const float* x, const float* t, int64_t n, int epochs, int batch, float lr,
                   uint64_t seed, float* final_loss) {
    if (!m || !x || !t || n < 1 || batch < 1) return set_error(AFFT_ERR_INVALID_ARGUMENT, "bad argument");
    AFFT_HIP(hipSetDevice(m->device));
    int rc = m->ensure(std::max<int64_t>(n, batch));
    if (rc) return rc;
    std::vector<float> xn, tn(n);
    m->normalise(x, n, xn);
    for (int64_t i = 0; i < n; ++i) tn[i] = (t[i] - m->ymean) / m->ystd;
    std::vector<int64_t> idx(n);
    for (int64_t i = 0; i < n; ++i) idx[i] = i;
    std::vector<float> bx((size_t)batch * MI), bt(batch);
    uint64_t c = seed * 0xD1342543DE82EF95ULL + 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < epochs; ++ep) {
        for (int64_t i = n - 1; i > 0; --i) std::swap(idx[i], idx[(int64_t)(sm64(c++) % (uint64_t)(i + 1))]);
        for (int64_t s0 = 0; s0 < n; s0 += batch) {
            const int64_t nb = std::min<int64_t>(batch, n - s0);
            for (int64_t k = 0; k < nb; ++k) {
                std::memcpy(&bx[k * MI], &xn[idx[s0 + k] * MI], MI * 4);
                bt[k] = tn[idx[s0 + k]];
            }
            AFFT_HIP(hipMemcpy(m->X, bx.data(), nb * MI * 4, hipMemcpyHostToDevice));
            AFFT_HIP(hipMemcpy(m->T, bt.data(), nb * 4, hipMemcpyHostToDevice));
            if ((rc = m->forward(nb)) || (rc = m->backward(nb))) return rc;
            ++m->step;
            float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
            void* a[] = {&m->P, &m->G, &m->Mm, &m->Vv, &lr, (void*)&b1, (void*)&b2, (void*)&eps, &c1, &c2};
            AFFT_HIP(hipLaunchKernel((const void*)&k_adam, dim3((NPARAM + 255) / 256), dim3(256), a, 0, nullptr));
        }
    }
    AFFT_HIP(hipDeviceSynchronize());
    if (final_loss) {
        std::vector<float> y(n);
        AFFT_HIP(hipMemcpy(m->X, xn.data(), xn.size() * 4, hipMemcpyHostToDevice));
        if ((rc = m->forward(n))) return rc;
        AFFT_HIP(hipMemcpy(y.data(), m->Y, n * 4, hipMemcpyDeviceToHost));
        double l = 0;
        for (int64_t i = 0; i < n; ++i) l += (double)(y[i] - tn[i]) * (y[i] - tn[i]);
        *final_loss = (float)(l / n);
    }
    clear_error();
    return AFFT_OK;
}

// Predicted runtime in the target's units (log ms as trained).
int afft_mlp_predict(afft_mlp m, const float* x, int64_t n, float* y) {
    if (!m || !x || !y || n < 1) return set_error(AFFT_ERR_INVALID_ARGUMENT, "bad argument");
    int rc = mlp_predict_host(m, x, n, y);
    if (rc == AFFT_OK) clear_error();
    return rc;
}

// Weights file: magic, version, parameter count, normalisation, parameters,
// FNV-1a checksum of everything before it.
int afft_mlp_save(afft_mlp m, const char* path) {
    if (!m || !path) return set_error(AFFT_ERR_NIL_SLICE, "NULL argument");
    std::vector<float> p(NPARAM);
    int rc = afft_mlp_get_params(m, p.data());
    if (rc) return rc;
    std::vector<unsigned char> buf;
    auto put = [&](const void* d, size_t k) {
        const unsigned char* c = (const unsigned char*)d;
        buf.insert(buf.end(), c, c + k);
    };
    const char magic[8] = {'A', 'F', 'F', 'T', 'M', 'L', 'P', '1'};
    const int32_t ver = 1, np = NPARAM, nf = NFEAT;
    put(magic, 8);
    put(&ver, 4);
    put(&np, 4);
    put(&nf, 4);
    put(m->mean, sizeof(m->mean));
    put(m->stdv, sizeof(m->stdv));
    put(&m->ymean, 4);
    put(&m->ystd, 4);
    put(p.data(), p.size() * 4);
    uint64_t h = 1469598103934665603ULL;
    for (unsigned char c : buf) h = (h ^ c) * 1099511628211ULL;
    put(&h, 8);
    FILE* f = std::fopen(path, "wb");
    if (!f) return set_error(AFFT_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path);
    const size_t w = std::fwrite(buf.data(), 1, buf.size(), f);
    std::fclose(f);
    if (w != buf.size()) return set_error(AFFT_ERR_INVALID_ARGUMENT, "short write");
    clear_error();
    return AFFT_OK;
}

int afft_mlp_load(afft_mlp m, const char* path) {
    if (!m || !path) return set_error(AFFT_ERR_NIL_SLICE, "NULL argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return set_error(AFFT_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path);
    std::vector<unsigned char> buf;
    unsigned char tmp[4096];
    size_t k;
    while ((k = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + k);
    std::fclose(f);
    const size_t body = 8 + 12 + 2 * NFEAT * 4 + 8 + (size_t)NPARAM * 4;
    if (buf.size() != body + 8 || std::memcmp(buf.data(), "AFFTMLP1", 8) != 0)
        return set_error(AFFT_ERR_INVALID_ARGUMENT, "not a cost-model weights file");
    uint64_t h = 1469598103934665603ULL, stored;
    for (size_t i = 0; i < body; ++i) h = (h ^ buf[i]) * 1099511628211ULL;
    std::memcpy(&stored, buf.data() + body, 8);
    if (h != stored) return set_error(AFFT_ERR_INVALID_ARGUMENT, "weights file checksum mismatch");
    int32_t np, nf;
    std::memcpy(&np, buf.data() + 12, 4);
    std::memcpy(&nf, buf.data() + 16, 4);
    if (np != NPARAM || nf != NFEAT) return set_error(AFFT_ERR_INVALID_ARGUMENT, "weights file shape mismatch");
    size_t o = 20;
    std::memcpy(m->mean, buf.data() + o, sizeof(m->mean));
    o += sizeof(m->mean);
    std::memcpy(m->stdv, buf.data() + o, sizeof(m->stdv));
    o += sizeof(m->stdv);
    std::memcpy(&m->ymean, buf.data() + o, 4);
    std::memcpy(&m->ystd, buf.data() + o + 4, 4);
    o += 8;
    return afft_mlp_set_params(m, (const float*)(buf.data() + o));
}

int afft_mlp_destroy(afft_mlp m) {
    delete m;
    clear_error();
    return AFFT_OK;
}

// Plans built with AFFT_STRATEGY_AUTO on this context consult the model when
// no wisdom entry matches (NULL detaches it).  The model must outlive its use.
int afft_context_set_cost_model(afft_context ctx, afft_mlp m) {
    if (!ctx) return set_error(AFFT_ERR_INVALID_ARGUMENT, "NULL context");
    std::lock_guard<std::mutex> l(ctx->mu);
    ctx->model = m;
    clear_error();
    return AFFT_OK;
}

}  // extern "C"
