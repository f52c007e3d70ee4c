// MFMA form of the radix-16 Stockham passes (AFFT_FLAG_MFMA, complex64).
//
// A radix-16 DFT of one butterfly is a 16x16 complex matrix product; its real
// embedding is 32x32.  With inputs interleaved as k = 2n + c (c = 0 real,
// 1 imaginary) and outputs as o = 2k' + c', gfx950's v_mfma_f32_16x16x32_f16
// (K = 32) computes outputs 0..15 of 16 butterflies in one instruction and
// outputs 16..31 in a second.  Operand lanes (wave of 64, q = lane/16,
// col = lane%16):
//   A: row col, columns 8q..8q+7        (the DFT matrix, fixed per lane)
//   B: rows 8q..8q+7, column col        (4 complex inputs r = 4q..4q+3 of
//                                        butterfly col, as (re, im) pairs)
//   D: rows 4q..4q+3, column col        (outputs r = 2q, 2q+1 of the first
//                                        product, 8+2q, 9+2q of the second)
// FP16 keeps 11 significant bits, so data and matrix are split into hi + lo
// halves and three products hi*hi + hi*lo + lo*hi accumulate in FP32
// (about 22 bits).  Before the split each butterfly is scaled by a power of
// two so its largest component lies in [0.5, 1): FP16 never overflows, and
// the scale is undone exactly on the outputs.  Twiddles are applied in FP32
// before the split.
//
// Lengths 256 (16x16), 1024 (4x16x16, the radix-4 pass on the vector ALUs)
// and 4096 (16x16x16).  A block holds G transforms in one LDS buffer (rows of
// pitch N + N/16, one pad slot per 16 points) and has one thread per 16
// points; each wave runs four 16-butterfly groups per pass, reading all of
// them before a barrier and writing after it, so the passes work in place.
// Rows (VAR_ROW, and the input of VAR_4B) are read and written straight from
// global memory, whose index runs along the butterflies; the four-step column
// pass (VAR_4A) and the transposed store of VAR_4B go through LDS staging
// loops whose lanes run along the transform index.
#include "afft_internal.hpp"

namespace afft {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int mf_rad(int N, int p) { return (N == 1024 && p == 0) ? 4 : 16; }
constexpr int mf_npass(int N) { return N == 256 ? 2 : 3; }
constexpr int mf_ns(int N, int p) {
    int ns = 1;
    for (int i = 0; i < p; ++i) ns *= mf_rad(N, i);
    return ns;
}
constexpr int mf_twoff(int N, int p) {
    int off = 0;
    for (int i = 1; i < p; ++i) off += mf_ns(N, i) * (mf_rad(N, i) - 1);
    return off;
}
__device__ __forceinline__ int mf_pad(int i) { return i + (i >> 4); }

// DFT-16 coefficient of embedded input kk for embedded output o.
template <int DIR>
__device__ __forceinline__ float dft16_coef(int o, int kk) {
    const int kp = o >> 1, cp = o & 1, n = kk >> 1, c = kk & 1;
    const int m = (n * kp) & 15;
    constexpr RootTable<16> tab{};
    const float cs = (float)tab.c[m], sn = DIR * (float)tab.s[m];
    if (cp == 0) return c == 0 ? cs : -sn;   // Re y = Re x cos - Im x (DIR sin)
    return c == 0 ? sn : cs;                  // Im y = Re x (DIR sin) + Im x cos
}

// 2^e as a float, e in [-126, 127]
__device__ __forceinline__ float pow2f(int e) { return __int_as_float((127 + e) << 23); }

// One radix-16 butterfly per lane column: x[i] holds input r = 4q + i; y gets
// outputs 2q, 2q+1, 8+2q, 9+2q.
__device__ __forceinline__ void mf_dft16(const half8 (&Ah)[2], const half8 (&Al)[2], const cx<float> (&x)[4],
                                         cx<float> (&y)[4]) {
    float m = 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) m = fmaxf(m, fmaxf(fabsf(x[i].x), fabsf(x[i].y)));
    m = fmaxf(m, __shfl_xor(m, 16));
    m = fmaxf(m, __shfl_xor(m, 32));
    // m = f * 2^e with f in [0.5, 1); e clamped so both scale factors exist
    int e = ((__float_as_int(m) >> 23) & 0xff) - 126;
    e = e > 126 ? 126 : e;
    const float down = pow2f(-e), up = pow2f(e);
    half8 bh, bl;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float re = x[i].x * down, im = x[i].y * down;
        const _Float16 hr = (_Float16)re, hi = (_Float16)im;
        bh[2 * i] = hr;
        bh[2 * i + 1] = hi;
        bl[2 * i] = (_Float16)(re - (float)hr);
        bl[2 * i + 1] = (_Float16)(im - (float)hi);
    }
    f32x4 d0 = {0.f, 0.f, 0.f, 0.f}, d1 = {0.f, 0.f, 0.f, 0.f};
    d0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[0], bh, d0, 0, 0, 0);
    d0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[0], bl, d0, 0, 0, 0);
    d0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Al[0], bh, d0, 0, 0, 0);
    d1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[1], bh, d1, 0, 0, 0);
    d1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[1], bl, d1, 0, 0, 0);
    d1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Al[1], bh, d1, 0, 0, 0);
    y[0] = mk<float>(d0[0] * up, d0[1] * up);
    y[1] = mk<float>(d0[2] * up, d0[3] * up);
    y[2] = mk<float>(d1[0] * up, d1[1] * up);
    y[3] = mk<float>(d1[2] * up, d1[3] * up);
}

template <int N, int G>
struct MfCfg {
    static constexpr int T = G * N / 16;          // threads (one per 16 points)
    static constexpr int WAVES = T / 64;
    static constexpr int NB = N / 16;             // radix-16 butterflies per transform
    static constexpr int GROUPS = NB / 16;        // 16-butterfly groups per transform
    static constexpr int PITCH = N + N / 16;
    static constexpr int P = mf_npass(N);
    static_assert(G * GROUPS == 4 * WAVES, "four groups per wave");
};

template <int N, int G, int VAR, int DIR, int p>
__device__ __forceinline__ void mf_pass(const KArgs& a, cx<float>* lds, const half8 (&Ah)[2], const half8 (&Al)[2],
                                        const cx<float>* __restrict__ in, cx<float>* __restrict__ out, long long c0) {
    using C = MfCfg<N, G>;
    constexpr int R = mf_rad(N, p), NS = mf_ns(N, p), TWO = mf_twoff(N, p);
    constexpr bool GIN = p == 0 && VAR != VAR_4A;          // first pass reads global rows
    constexpr bool GOUT = p == C::P - 1 && VAR == VAR_ROW; // last pass writes global rows
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int q = lane >> 4, col = lane & 15;
    const cx<float>* __restrict__ tw = (const cx<float>*)a.tw;
    if constexpr (R == 16) {
        cx<float> y[4][4];
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const int item = wave + C::WAVES * it;
            const int c = item / C::GROUPS, g = item - c * C::GROUPS;
            const int jb = 16 * g + col;
            const long long cg = c0 + c;
            const bool act = cg < a.count;
            cx<float> x[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int r = 4 * q + i;
                const int n = jb + r * C::NB;
                if constexpr (GIN) x[i] = act ? in[cg * a.in_cd + n] : mk<float>(0, 0);
                else x[i] = lds[c * C::PITCH + mf_pad(n)];
                if constexpr (NS > 1) {
                    if (r > 0) x[i] = twmul<DIR>(x[i], tw[TWO + (r - 1) * NS + jb % NS]);   // r-major table
                }
            }
            mf_dft16(Ah, Al, x, y[it]);
        }
        if constexpr (!GIN) __syncthreads();   // every read of this pass precedes any write
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const int item = wave + C::WAVES * it;
            const int c = item / C::GROUPS, g = item - c * C::GROUPS;
            const int jb = 16 * g + col;
            const int base = (jb / NS) * NS * 16 + jb % NS;
            const int ko[4] = {2 * q, 2 * q + 1, 8 + 2 * q, 9 + 2 * q};
            const long long cg = c0 + c;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int idx = base + ko[i] * NS;
                if constexpr (GOUT) {
                    if (cg < a.count) st_out(out + cg * a.out_cd + idx, scl(y[it][i], (float)a.scale), a.nt);
                } else {
                    lds[c * C::PITCH + mf_pad(idx)] = y[it][i];
                }
            }
        }
    } else {
        // radix 4 on the vector ALUs (first pass of N = 1024, Ns = 1):
        // butterflies b = tid + T*i of the block, transform b / 256
        static_assert(R == 4 && p == 0 && NS == 1, "radix 4 only as the first pass");
        cx<float> v[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int b = tid + C::T * i;
            const int c = b / (N / 4), j = b - c * (N / 4);
            const long long cg = c0 + c;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int n = j + r * (N / 4);
                if constexpr (GIN) v[i][r] = cg < a.count ? in[cg * a.in_cd + n] : mk<float>(0, 0);
                else v[i][r] = lds[c * C::PITCH + mf_pad(n)];
            }
            dft<4, DIR, 1>(v[i]);
        }
        if constexpr (!GIN) __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int b = tid + C::T * i;
            const int c = b / (N / 4), j = b - c * (N / 4);
#pragma unroll
            for (int r = 0; r < 4; ++r) lds[c * C::PITCH + mf_pad(4 * j + r)] = v[i][r];
        }
    }
    if constexpr (!GOUT) __syncthreads();
}

template <int N, int G, int VAR, int DIR, int... Ps>
__device__ __forceinline__ void mf_passes(const KArgs& a, cx<float>* lds, const half8 (&Ah)[2], const half8 (&Al)[2],
                                          const cx<float>* in, cx<float>* out, long long c0,
                                          std::integer_sequence<int, Ps...>) {
    (mf_pass<N, G, VAR, DIR, Ps>(a, lds, Ah, Al, in, out, c0), ...);
}

template <int N, int G, int VAR, int DIR>
__global__ void __launch_bounds__((MfCfg<N, G>::T)) k_mfma(KArgs a) {
    using C = MfCfg<N, G>;
    __shared__ cx<float> lds[G * C::PITCH];
    const int tid = threadIdx.x, lane = tid & 63;
    const int q = lane >> 4, col = lane & 15;
    const long long c0 = (long long)blockIdx.x * G;
    const cx<float>* __restrict__ in = (const cx<float>*)a.in + (long long)blockIdx.y * a.in_bd;
    cx<float>* __restrict__ out = (cx<float>*)a.out + (long long)blockIdx.y * a.out_bd;

    // A operands (fixed): [product][hi/lo]
    half8 Ah[2], Al[2];
#pragma unroll
    for (int pr = 0; pr < 2; ++pr)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float v = dft16_coef<DIR>(16 * pr + col, 8 * q + e);
            const _Float16 h = (_Float16)v;
            Ah[pr][e] = h;
            Al[pr][e] = (_Float16)(v - (float)h);
        }

    if constexpr (VAR == VAR_4A) {
        // columns: lanes run along the transform index c
        for (int i = tid; i < G * N; i += C::T) {
            const int n = i / G, c = i - n * G;
            const long long cg = c0 + c;
            lds[c * C::PITCH + mf_pad(n)] = cg < a.count ? in[cg * a.in_cd + (long long)n * a.in_nd] : mk<float>(0, 0);
        }
        __syncthreads();
    }
    mf_passes<N, G, VAR, DIR>(a, lds, Ah, Al, in, out, c0, std::make_integer_sequence<int, C::P>{});
    if constexpr (VAR == VAR_4A || VAR == VAR_4B) {
        // thread tid stores points k = tid/G + s*(T/G) of column c = tid mod G
        constexpr int STEPS = G * N / C::T, KS = C::T / G;
        static_assert(C::T % G == 0 && (G * N) % C::T == 0, "whole columns per step");
        const float sc = (float)a.scale;
        const int c = tid % G, k0 = tid / G;
        const long long cg = c0 + c;
        const long long col = cg + a.tw4_c0;
        // W_L^{k*col}: a table value every fourth step, the step factor
        // W_L^{KS*col} in between (the table gathers are scattered lines)
        cx<float> st = mk<float>(1, 0), w = mk<float>(1, 0);
        if constexpr (VAR == VAR_4A) st = tw4<float>(a, (long long)KS * col);
#pragma unroll
        for (int s = 0; s < STEPS; ++s) {
            const int k = k0 + s * KS;
            if constexpr (VAR == VAR_4A) w = (s % 4 == 0) ? tw4<float>(a, (long long)k * col) : cmul(w, st);
            if (cg < a.count) {
                cx<float> v = lds[c * C::PITCH + mf_pad(k)];
                if constexpr (VAR == VAR_4A) v = twmul<DIR>(v, w);
                st_out(out + cg * a.out_cd + (long long)k * a.out_nd, scl(v, sc), a.nt);
            }
        }
    }
}

// Kernel, transforms per block and threads per block of the MFMA form for a
// length and layout; nullptr when the form does not cover them.
const void* mfma_kernel(int n, int var, int dir, int* per_block, int* threads) {
#define AFFT_MF(NN, GG, VV)                                                                  \
    if (n == NN && var == VV) {                                                              \
        *per_block = GG;                                                                     \
        *threads = MfCfg<NN, GG>::T;                                                         \
        return dir < 0 ? (const void*)&k_mfma<NN, GG, VV, -1> : (const void*)&k_mfma<NN, GG, VV, 1>; \
    }
    AFFT_MF(256, 16, VAR_ROW)
    AFFT_MF(1024, 4, VAR_ROW)
    AFFT_MF(4096, 1, VAR_ROW)
    AFFT_MF(256, 16, VAR_4A)
    AFFT_MF(1024, 16, VAR_4A)
    AFFT_MF(256, 16, VAR_4B)
    AFFT_MF(1024, 16, VAR_4B)
#undef AFFT_MF
    return nullptr;
}

}  // namespace afft
