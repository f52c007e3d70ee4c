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
// (about 22 bits).  Twiddles are applied in FP32 before the split.
//
// Scaling.  Each transform is scaled once, by a power of two 2^-e chosen so
// that its largest input component lies in [0.5, 1); the scale is undone
// exactly on the final store.  A radix-R pass grows the largest complex
// modulus by at most R, so the inputs of the MFMA passes of a transform of
// up to 16384 points stay below sqrt(2) * 4096 < 65504 and FP16 never
// overflows; the hi/lo split keeps 22 bits down to components of 2^-14 of
// the transform's largest input and an absolute 2^-25 of it below that
// (FP16 subnormals).  The per-transform maximum is one wave reduction per
// transform and wave plus an LDS atomic, once per kernel.
//
// Lengths 256 .. 16384: radix-16 passes on the matrix cores, with a first
// radix-2/4/8 pass on the vector ALUs when the length is not a power of 16.
// A block holds G transforms in one LDS buffer (rows of pitch N + N/16, one
// pad slot per 16 points) and has one thread per 16 points; each wave runs
// four 16-butterfly groups per pass, reading all of them before a barrier and
// writing after it, so the passes work in place.  Rows (VAR_ROW, and the
// input of VAR_4B) are read and written straight from global memory, whose
// index runs along the butterflies; the four-step column pass (VAR_4A) and
// the transposed store of VAR_4B go through LDS staging loops whose lanes run
// along the transform index.
#include "afft_internal.hpp"

namespace afft {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int mf_lg(int N) { return N <= 1 ? 0 : 1 + mf_lg(N / 2); }
// remainder radix of the first pass (1: none, the length is a power of 16)
constexpr int mf_r0(int N) { return 1 << (mf_lg(N) % 4); }
constexpr int mf_npass(int N) { return mf_lg(N) / 4 + (mf_r0(N) > 1 ? 1 : 0); }
constexpr int mf_rad(int N, int p) { return (p == 0 && mf_r0(N) > 1) ? mf_r0(N) : 16; }
// the first pass on the matrix cores
constexpr int mf_p0(int N) { return mf_r0(N) > 1 ? 1 : 0; }
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

// Exponent e of the transform scale 2^-e from the bits of its largest input
// component m (m = f * 2^e with f in [0.5, 1)), clamped so that both 2^-e
// and 2^e are normal floats.
__device__ __forceinline__ int mf_exp(unsigned mbits) {
    int e = (int)((mbits >> 23) & 0xff) - 126;
    return e > 126 ? 126 : (e < -126 ? -126 : e);
}

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, 0xf, 0xf, true));
}

// Maximum over the lanes l ^ x of every xor mask x >= XMIN (a power of two):
// the whole wave for XMIN = 1.  Within rows of 16 lanes by DPP row
// rotations (after rotating by 8, 4, 2, 1 lane i holds the maximum over the
// lanes i + k*XMIN of its row), across rows by lane shuffles.
template <int XMIN>
__device__ __forceinline__ float wave_max(float v) {
    if constexpr (XMIN <= 8) v = fmaxf(v, dpp_f<0x128>(v));   // row_ror:8
    if constexpr (XMIN <= 4) v = fmaxf(v, dpp_f<0x124>(v));   // row_ror:4
    if constexpr (XMIN <= 2) v = fmaxf(v, dpp_f<0x122>(v));   // row_ror:2
    if constexpr (XMIN <= 1) v = fmaxf(v, dpp_f<0x121>(v));   // row_ror:1
    if constexpr (XMIN <= 16) v = fmaxf(v, __shfl_xor(v, 16));
    v = fmaxf(v, __shfl_xor(v, 32));
    return v;
}

__device__ __forceinline__ float cmax(cx<float> v, float m) { return fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y))); }

// One radix-16 butterfly per lane column: x[i] holds input r = 4q + i
// (already scaled); y gets outputs 2q, 2q+1, 8+2q, 9+2q.  y may be x: every
// input is read before the first output is written.
__device__ __forceinline__ void mf_dft16(const half8 (&Ah)[2], const half8 (&Al)[2], const cx<float> (&x)[4],
                                         cx<float> (&y)[4]) {
    half8 bh, bl;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float re = x[i].x, im = x[i].y;
        const _Float16 hr = (_Float16)re, hi = (_Float16)im;
        bh[2 * i] = hr;
        bh[2 * i + 1] = hi;
        // lo = x - hi, exact in FP32 and rounded once (one mixed-precision FMA)
        bl[2 * i] = (_Float16)__builtin_fmaf(re, 1.0f, -(float)hr);
        bl[2 * i + 1] = (_Float16)__builtin_fmaf(im, 1.0f, -(float)hi);
    }
    f32x4 d0 = {0.f, 0.f, 0.f, 0.f}, d1 = {0.f, 0.f, 0.f, 0.f};
    d0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[0], bh, d0, 0, 0, 0);
    d0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[0], bl, d0, 0, 0, 0);
    d0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Al[0], bh, d0, 0, 0, 0);
    d1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[1], bh, d1, 0, 0, 0);
    d1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ah[1], bl, d1, 0, 0, 0);
    d1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Al[1], bh, d1, 0, 0, 0);
    y[0] = mk<float>(d0[0], d0[1]);
    y[1] = mk<float>(d0[2], d0[3]);
    y[2] = mk<float>(d1[0], d1[1]);
    y[3] = mk<float>(d1[2], d1[3]);
}

template <int N, int G>
struct MfCfg {
    static constexpr int T = G * N / 16;          // threads (one per 16 points)
    static constexpr int WAVES = T / 64;
    static constexpr int NB = N / 16;             // radix-16 butterflies per transform
    static constexpr int GROUPS = NB / 16;        // 16-butterfly groups per transform
    static constexpr int PITCH = N + N / 16;
    static constexpr int P = mf_npass(N);
    static constexpr int P0 = mf_p0(N);
    static_assert(G * GROUPS == 4 * WAVES, "four groups per wave");
    static_assert(T % 64 == 0 && T <= 1024, "whole waves, at most 1024 threads");
};

// Twiddles of radix-16 pass p (Ns > 1) for the four items of this lane:
// w^r of butterfly jb % Ns for r = 4q + i, from the r-major table (entry
// (r-1)*Ns + jb % Ns; r = 0, for q = 0 and i = 0, takes entry r = 1's slot
// and is discarded, so no lane branches around a load).
template <int N, int G, int p>
__device__ __forceinline__ void mf_tw_load(const cx<float>* __restrict__ tw, cx<float> (&wt)[4][4]) {
    using C = MfCfg<N, G>;
    constexpr int NS = mf_ns(N, p), TWO = mf_twoff(N, p);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int q = lane >> 4, col = lane & 15;
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        const int item = wave + C::WAVES * it;
        const int c = item / C::GROUPS, g = item - c * C::GROUPS;
        const int jb = 16 * g + col;
        const int tq = TWO + jb % NS + (4 * q - 1) * NS;
#pragma unroll
        for (int i = 0; i < 4; ++i) wt[it][i] = tw[i > 0 ? tq + i * NS : (q > 0 ? tq : tq + NS)];
    }
}

// Pass p.  Its twiddles arrive in wt, loaded by pass p - 1 before that
// pass's closing barrier, so their latency overlaps the barrier wait; it
// loads pass p + 1's the same way.  (Loaded where they are used, the
// dependent global loads cost the N=4096 kernel about 20 % of its time.)
template <int N, int G, int VAR, int DIR, int p>
__device__ __forceinline__ void mf_pass(const KArgs& a, cx<float>* lds, unsigned* smax, const half8 (&Ah)[2],
                                        const half8 (&Al)[2], const cx<float>* __restrict__ in,
                                        cx<float>* __restrict__ out, long long c0, cx<float> (&wt)[4][4]) {
    using C = MfCfg<N, G>;
    constexpr int R = mf_rad(N, p), NS = mf_ns(N, p);
    constexpr bool GIN = p == 0 && VAR != VAR_4A;          // first pass reads global rows
    constexpr bool GOUT = p == C::P - 1 && VAR == VAR_ROW; // last pass writes global rows
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int q = lane >> 4, col = lane & 15;
    const cx<float>* __restrict__ tw = (const cx<float>*)a.tw;
    if constexpr (R == 16) {
        cx<float> x[4][4];   // inputs, then (in place) outputs of the four items
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const int item = wave + C::WAVES * it;
            const int c = item / C::GROUPS, g = item - c * C::GROUPS;
            const int jb = 16 * g + col;
            const long long cg = c0 + c;
            const bool act = cg < a.count;
            // rows past the batch read the last row (no branch) and are zeroed
            const long long cl = act ? cg : a.count - 1;
            // inputs n = n0 + i*NB; NB is a multiple of 16, so the padded LDS
            // index of input i is the lane's base plus a constant
            const int n0 = jb + 4 * q * C::NB;
            const cx<float>* __restrict__ rp = in + (cl * a.in_cd + n0);
            const int lb = c * C::PITCH + mf_pad(n0);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if constexpr (GIN) {
                    const cx<float> v = rp[i * C::NB];
                    x[it][i] = act ? v : mk<float>(0, 0);
                } else {
                    x[it][i] = lds[lb + i * (C::NB + C::NB / 16)];
                }
            }
        }
        if constexpr (GIN) {
            // this pass reads the input: the per-transform maximum (every
            // lane of an item belongs to the item's transform; with one
            // transform per block, the four items of a wave share it)
            if constexpr (G == 1) {
                float m = 0.f;
#pragma unroll
                for (int it = 0; it < 4; ++it)
#pragma unroll
                    for (int i = 0; i < 4; ++i) m = cmax(x[it][i], m);
                m = wave_max<1>(m);
                if (lane == 0) atomicMax(&smax[0], __float_as_uint(m));
            } else {
#pragma unroll
                for (int it = 0; it < 4; ++it) {
                    float m = 0.f;
#pragma unroll
                    for (int i = 0; i < 4; ++i) m = cmax(x[it][i], m);
                    m = wave_max<1>(m);
                    const int item = wave + C::WAVES * it;
                    if (lane == 0) atomicMax(&smax[item / C::GROUPS], __float_as_uint(m));
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const int item = wave + C::WAVES * it;
            const int c = item / C::GROUPS;
            float down = 1.f;
            if constexpr (p == C::P0) down = pow2f(-mf_exp(smax[c]));
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int r = 4 * q + i;
                if constexpr (NS > 1) x[it][i] = r > 0 ? twmul<DIR>(x[it][i], wt[it][i]) : x[it][i];
                if constexpr (p == C::P0) x[it][i] = scl(x[it][i], down);
            }
            mf_dft16(Ah, Al, x[it], x[it]);
        }
        if constexpr (!GIN) __syncthreads();   // every read of this pass precedes any write
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const int item = wave + C::WAVES * it;
            const int c = item / C::GROUPS, g = item - c * C::GROUPS;
            const int jb = 16 * g + col;
            // outputs idx = base + 2q*NS + {0, 1, 8, 9}*NS
            const int b2 = (jb / NS) * NS * 16 + jb % NS + 2 * q * NS;
            constexpr int KO[4] = {0, 1, 8, 9};
            const long long cg = c0 + c;
            float sc = 1.f;
            if constexpr (GOUT) sc = (float)a.scale * pow2f(mf_exp(smax[c]));
            cx<float>* __restrict__ op = out + (cg * a.out_cd + b2);
            // padded LDS index: with NS a multiple of 16 the padding of the
            // constant part separates; with NS = 1 the outputs 16 jb + ko
            // (ko < 16) pad to 17 jb + ko
            const int lb = c * C::PITCH + (NS % 16 == 0 ? mf_pad(b2) : NS == 1 ? 17 * jb + 2 * q : 0);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if constexpr (GOUT) {
                    if (cg < a.count) st_out(op + KO[i] * NS, scl(x[it][i], sc), a.nt);
                } else if constexpr (NS % 16 == 0) {
                    lds[lb + KO[i] * (NS + NS / 16)] = x[it][i];
                } else if constexpr (NS == 1) {
                    lds[lb + KO[i]] = x[it][i];
                } else {
                    lds[c * C::PITCH + mf_pad(b2 + KO[i] * NS)] = x[it][i];
                }
            }
        }
    } else {
        // radix 2, 4 or 8 on the vector ALUs (first pass, Ns = 1):
        // butterflies b = tid + T*i of the block, transform b / (N/R)
        static_assert(p == 0 && NS == 1 && 16 % R == 0, "small radix only as the first pass");
        constexpr int NBUT = 16 / R;
        cx<float> v[NBUT][R];
#pragma unroll
        for (int i = 0; i < NBUT; ++i) {
            const int b = tid + C::T * i;
            const int c = b / (N / R), j = b - c * (N / R);
            const long long cg = c0 + c;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int n = j + r * (N / R);
                if constexpr (GIN) v[i][r] = cg < a.count ? in[cg * a.in_cd + n] : mk<float>(0, 0);
                else v[i][r] = lds[c * C::PITCH + mf_pad(n)];
            }
            if constexpr (GIN) {
                // the per-transform maximum of the input (a wave's butterflies
                // i belong to one transform: N/R is a multiple of 64)
                float m = 0.f;
#pragma unroll
                for (int r = 0; r < R; ++r) m = cmax(v[i][r], m);
                m = wave_max<1>(m);
                if (lane == 0) atomicMax(&smax[c], __float_as_uint(m));
            }
            dft<R, DIR, 1>(v[i]);
        }
        if constexpr (!GIN) __syncthreads();
#pragma unroll
        for (int i = 0; i < NBUT; ++i) {
            const int b = tid + C::T * i;
            const int c = b / (N / R), j = b - c * (N / R);
#pragma unroll
            for (int r = 0; r < R; ++r) lds[c * C::PITCH + mf_pad(R * j + r)] = v[i][r];
        }
    }
    if constexpr (p + 1 < C::P) {
        if constexpr (mf_rad(N, p + 1) == 16 && mf_ns(N, p + 1) > 1) mf_tw_load<N, G, p + 1>(tw, wt);
    }
    if constexpr (!GOUT) __syncthreads();
}

template <int N, int G, int VAR, int DIR, int... Ps>
__device__ __forceinline__ void mf_passes(const KArgs& a, cx<float>* lds, unsigned* smax, const half8 (&Ah)[2],
                                          const half8 (&Al)[2], const cx<float>* in, cx<float>* out, long long c0,
                                          std::integer_sequence<int, Ps...>) {
    cx<float> wt[4][4];
    (mf_pass<N, G, VAR, DIR, Ps>(a, lds, smax, Ah, Al, in, out, c0, wt), ...);
}

template <int N, int G, int VAR, int DIR>
__global__ void __launch_bounds__((MfCfg<N, G>::T)) k_mfma(KArgs a) {
    using C = MfCfg<N, G>;
    __shared__ cx<float> lds[G * C::PITCH];
    __shared__ unsigned smax[G];   // per-transform largest input component (float bits)
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int q = lane >> 4, col = lane & 15;
    const long long c0 = (long long)blockIdx.x * G;
    const cx<float>* __restrict__ in = (const cx<float>*)a.in + (long long)blockIdx.y * a.in_bd;
    cx<float>* __restrict__ out = (cx<float>*)a.out + (long long)blockIdx.y * a.out_bd;
    if (tid < G) smax[tid] = 0u;

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
    __syncthreads();   // smax cleared

    if constexpr (VAR == VAR_4A) {
        // columns: lanes run along the transform index c = tid mod G, which
        // is the same in every step
        static_assert(C::T % G == 0, "one column per thread");
        float m = 0.f;
        for (int i = tid; i < G * N; i += C::T) {
            const int n = i / G, c = i - n * G;
            const long long cg = c0 + c;
            const cx<float> v = cg < a.count ? in[cg * a.in_cd + (long long)n * a.in_nd] : mk<float>(0, 0);
            m = cmax(v, m);
            lds[c * C::PITCH + mf_pad(n)] = v;
        }
        m = wave_max<G>(m);   // over the lanes of this column
        if (lane < G) atomicMax(&smax[lane], __float_as_uint(m));
        __syncthreads();
    }
    mf_passes<N, G, VAR, DIR>(a, lds, smax, Ah, Al, in, out, c0, std::make_integer_sequence<int, C::P>{});
    if constexpr (VAR == VAR_4A || VAR == VAR_4B) {
        // thread tid stores points k = tid/G + s*(T/G) of column c = tid mod G
        constexpr int STEPS = G * N / C::T, KS = C::T / G;
        static_assert(C::T % G == 0 && (G * N) % C::T == 0, "whole columns per step");
        const int c = tid % G, k0 = tid / G;
        const float sc = (float)a.scale * pow2f(mf_exp(smax[c]));
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
// length and layout; nullptr when the form does not cover them.  Rows take
// G = max(1, 4096/N) transforms per block; the four-step passes G columns
// within 1024 threads and 128 KiB of LDS (16 up to 1024 points, then 8 and 4).
const void* mfma_kernel(int n, int var, int dir, int* per_block, int* threads) {
#define AFFT_MF(NN, GG, VV)                                                                  \
    if (n == NN && var == VV) {                                                              \
        *per_block = GG;                                                                     \
        *threads = MfCfg<NN, GG>::T;                                                         \
        return dir < 0 ? (const void*)&k_mfma<NN, GG, VV, -1> : (const void*)&k_mfma<NN, GG, VV, 1>; \
    }
    AFFT_MF(256, 16, VAR_ROW)
    AFFT_MF(512, 8, VAR_ROW)
    AFFT_MF(1024, 4, VAR_ROW)
    AFFT_MF(2048, 2, VAR_ROW)
    AFFT_MF(4096, 1, VAR_ROW)
    AFFT_MF(8192, 1, VAR_ROW)
    AFFT_MF(16384, 1, VAR_ROW)
    AFFT_MF(256, 16, VAR_4A)
    AFFT_MF(512, 16, VAR_4A)
    AFFT_MF(1024, 16, VAR_4A)
    AFFT_MF(2048, 8, VAR_4A)
    AFFT_MF(4096, 4, VAR_4A)
    AFFT_MF(256, 16, VAR_4B)
    AFFT_MF(512, 16, VAR_4B)
    AFFT_MF(1024, 16, VAR_4B)
    AFFT_MF(2048, 8, VAR_4B)
    AFFT_MF(4096, 4, VAR_4B)
#undef AFFT_MF
    return nullptr;
}

}  // namespace afft
