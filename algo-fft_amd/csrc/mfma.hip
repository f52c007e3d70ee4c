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
// before the split.  Passes ping-pong between two LDS buffers.
#include "afft_internal.hpp"

namespace afft {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// DFT-16 coefficient of embedded input kk for embedded output o.
template <int DIR>
__device__ __forceinline__ float dft16_coef(int o, int kk) {
    const int kp = o >> 1, cp = o & 1, n = kk >> 1, c = kk & 1;
    const int m = (n * kp) & 15;
    // cos / sin of 2*pi*m/16 from the compile-time tables
    constexpr float ct[16] = {
        (float)cos2pi(0, 16), (float)cos2pi(1, 16), (float)cos2pi(2, 16), (float)cos2pi(3, 16),
        (float)cos2pi(4, 16), (float)cos2pi(5, 16), (float)cos2pi(6, 16), (float)cos2pi(7, 16),
        (float)cos2pi(8, 16), (float)cos2pi(9, 16), (float)cos2pi(10, 16), (float)cos2pi(11, 16),
        (float)cos2pi(12, 16), (float)cos2pi(13, 16), (float)cos2pi(14, 16), (float)cos2pi(15, 16)};
    constexpr float st[16] = {
        (float)sin2pi(0, 16), (float)sin2pi(1, 16), (float)sin2pi(2, 16), (float)sin2pi(3, 16),
        (float)sin2pi(4, 16), (float)sin2pi(5, 16), (float)sin2pi(6, 16), (float)sin2pi(7, 16),
        (float)sin2pi(8, 16), (float)sin2pi(9, 16), (float)sin2pi(10, 16), (float)sin2pi(11, 16),
        (float)sin2pi(12, 16), (float)sin2pi(13, 16), (float)sin2pi(14, 16), (float)sin2pi(15, 16)};
    const float cs = ct[m], sn = DIR * st[m];
    if (cp == 0) return c == 0 ? cs : -sn;   // Re y = Re x cos - Im x (DIR sin)
    return c == 0 ? sn : cs;                  // Im y = Re x (DIR sin) + Im x cos
}

template <int N, int DIR>
__global__ void __launch_bounds__(256) k_mfma(KArgs a) {
    constexpr int NB = N / 16;                 // butterflies per pass
    constexpr int GROUPS = NB / 16;            // MFMA groups of 16 butterflies
    constexpr int WPT = GROUPS < 4 ? GROUPS : 4;   // waves per transform
    constexpr int TPB = 4 / WPT;               // transforms per block
    constexpr int P = N == 256 ? 2 : 3;
    static_assert(N == 256 || N == 4096, "MFMA form is built for N = 256 and 4096");
    __shared__ cx<float> lds[2][TPB][N];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int q = lane >> 4, col = lane & 15;
    const int c = wave / WPT, wt = wave % WPT;          // transform in block, wave within it
    const long long cg = (long long)blockIdx.x * TPB + c;
    const bool act = cg < a.count;
    const cx<float>* __restrict__ tw = (const cx<float>*)a.tw;

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

    // global -> LDS (natural order)
    {
        const cx<float>* src = (const cx<float>*)a.in + cg * a.in_cd;
        const int t = tid - c * 64 * WPT;
        for (int n = t; n < N; n += 64 * WPT) lds[0][c][n] = act ? src[n] : mk<float>(0, 0);
    }
    __syncthreads();

    int ns = 1;
    int twoff = 0;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const cx<float>* rd = lds[p & 1][c];
        cx<float>* wr = lds[(p + 1) & 1][c];
        for (int g = wt; g < GROUPS; g += WPT) {
            const int jb = 16 * g + col;
            const int k = jb % ns;
            cx<float> x[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int r = 4 * q + i;
                x[i] = rd[jb + r * NB];
                if (p > 0 && r > 0) x[i] = twmul<DIR>(x[i], tw[twoff + k * 15 + r - 1]);
            }
            // per-butterfly power-of-two scale from the max over its 4 lanes
            float m = 0.f;
#pragma unroll
            for (int i = 0; i < 4; ++i) m = fmaxf(m, fmaxf(fabsf(x[i].x), fabsf(x[i].y)));
            m = fmaxf(m, __shfl_xor(m, 16));
            m = fmaxf(m, __shfl_xor(m, 32));
            int ex = 0;
            if (m > 0.f) (void)frexpf(m, &ex);
            half8 bh, bl;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float re = ldexpf(x[i].x, -ex), im = ldexpf(x[i].y, -ex);
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
            const int base = (jb / ns) * ns * 16 + (jb % ns);
            wr[base + (2 * q) * ns] = mk<float>(ldexpf(d0[0], ex), ldexpf(d0[1], ex));
            wr[base + (2 * q + 1) * ns] = mk<float>(ldexpf(d0[2], ex), ldexpf(d0[3], ex));
            wr[base + (8 + 2 * q) * ns] = mk<float>(ldexpf(d1[0], ex), ldexpf(d1[1], ex));
            wr[base + (9 + 2 * q) * ns] = mk<float>(ldexpf(d1[2], ex), ldexpf(d1[3], ex));
        }
        if (p > 0) twoff += ns * 15;
        ns *= 16;
        __syncthreads();
    }
    if (act) {
        cx<float>* dst = (cx<float>*)a.out + cg * a.out_cd;
        const int t = tid - c * 64 * WPT;
        const float s = (float)a.scale;
        for (int n = t; n < N; n += 64 * WPT) dst[n] = scl(lds[P & 1][c][n], s);
    }
}

const void* mfma_kernel(int n, int dir, int* tpb) {
    if (n == 256) {
        *tpb = 4;
        return dir < 0 ? (const void*)&k_mfma<256, -1> : (const void*)&k_mfma<256, 1>;
    }
    if (n == 4096) {
        *tpb = 1;
        return dir < 0 ? (const void*)&k_mfma<4096, -1> : (const void*)&k_mfma<4096, 1>;
    }
    return nullptr;
}

}  // namespace afft
