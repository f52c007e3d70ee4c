// Batched LDS-resident FFT kernels built on the compile-time Stockham core.
//
// Addressing of transform c (blockIdx.x*TPB + local c) in slice blockIdx.y:
//   element n at  base + blockIdx.y*bd + c*cd + n*nd      (separately for in/out)
// Variants:
//   VAR_ROW  each transform contiguous (nd = 1): ForwardBatch, N-D last axis
//   VAR_COL  TPB adjacent columns (cd = 1), element stride nd: strided plans,
//            N-D outer axes
//   VAR_4A   VAR_COL + four-step twiddle W_L^{k*c} on the outputs (L = N*count)
//   VAR_4B   rows in (nd = 1), outputs scattered as columns (cd = 1) through an
//            LDS re-map so both sides stay coalesced: the transposing second
//            pass of the four-step transform
#pragma once
#include <utility>

#include "fft_core.hpp"

namespace afft {

enum { VAR_ROW = 0, VAR_COL = 1, VAR_4A = 2, VAR_4B = 3, VAR_COUNT = 4 };

struct KArgs {
    const void* in;
    void* out;
    const void* tw;          // per-pass twiddle table of the schedule
    long long count;         // transforms along c
    long long in_cd, out_cd;
    long long in_nd, out_nd;
    long long in_bd, out_bd;
    double scale;            // applied when != 1 (inverse 1/N)
    const void* tw4_lo;      // four-step: W_L^m = hi[m >> sh] * lo[m & (2^sh-1)]
    const void* tw4_hi;
    int tw4_shift;
    long long tw4_L;
    long long tw4_c0;        // column index offset of this launch
};

template <typename T, int N>
struct PadFor {
    static constexpr int value = N >= 32 ? (sizeof(T) == 4 ? 4 : 3) : 0;
};

template <typename T, class S, int E, int TPB, int VAR>
struct KCfg {
    static constexpr int N = S::N;
    static constexpr int TT = N / E;
    static constexpr int THREADS = TPB * TT;
    static constexpr bool ROWL = (VAR == VAR_ROW || VAR == VAR_4B);
    static constexpr int PADSH = PadFor<T, N>::value;
    static constexpr int LDS_ELEMS = S::P > 1 || VAR == VAR_4B
                                         ? (ROWL ? TPB * LdsRow<N, PADSH>::SLAB : TPB * N)
                                         : 1;
};

// Forward-sign W_L^m from the two-level table; callers conjugate for DIR=+1.
template <typename T>
__device__ __forceinline__ cx<T> tw4(const KArgs& a, long long m) {
    const cx<T>* lo = (const cx<T>*)a.tw4_lo;
    const cx<T>* hi = (const cx<T>*)a.tw4_hi;
    const long long mask = (1LL << a.tw4_shift) - 1;
    return cmul(lo[m & mask], hi[m >> a.tw4_shift]);
}

template <typename T, class S, int E, int TPB, int VAR, int DIR>
__global__ void __launch_bounds__((KCfg<T, S, E, TPB, VAR>::THREADS)) k_fft(KArgs a) {
    using C = KCfg<T, S, E, TPB, VAR>;
    constexpr int N = S::N;
    constexpr int TT = C::TT;
    __shared__ cx<T> lds[C::LDS_ELEMS];
    const int tid = threadIdx.x;
    int c, t;
    if constexpr (C::ROWL) { c = tid / TT; t = tid % TT; }
    else { c = tid % TPB; t = tid / TPB; }
    const long long cg = (long long)blockIdx.x * TPB + c;
    const bool act = cg < a.count;
    const cx<T>* __restrict__ in = (const cx<T>*)a.in + (long long)blockIdx.y * a.in_bd;
    const cx<T>* __restrict__ tw = (const cx<T>*)a.tw;
    cx<T> v[E];

    if constexpr (VAR == VAR_ROW || VAR == VAR_4B) {
        const cx<T>* src = in + cg * a.in_cd;
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = act ? src[t + TT * e] : mk<T>(0, 0);
    } else {
        const cx<T>* src = in + cg;
        const long long nd = a.in_nd;
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = act ? src[(long long)(t + TT * e) * nd] : mk<T>(0, 0);
    }

    if constexpr (C::ROWL) {
        LdsRow<N, C::PADSH> li(c);
        core<T, S, E, DIR>(v, lds, t, tw, li);
    } else {
        LdsCol<TPB> li(c);
        core<T, S, E, DIR>(v, lds, t, tw, li);
    }

    if constexpr (VAR == VAR_4A) {
        const long long col = cg + a.tw4_c0;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            // k < N and col < L/N, so k*col < L needs no reduction
            const long long m = (long long)(t + TT * e) * col;
            v[e] = twmul<DIR>(v[e], tw4<T>(a, m));
        }
    }
    if (a.scale != 1.0) {
        const T s = (T)a.scale;
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = scl(v[e], s);
    }

    cx<T>* __restrict__ out = (cx<T>*)a.out + (long long)blockIdx.y * a.out_bd;
    if constexpr (VAR == VAR_ROW) {
        if (act) {
            cx<T>* dst = out + cg * a.out_cd;
#pragma unroll
            for (int e = 0; e < E; ++e) dst[t + TT * e] = v[e];
        }
    } else if constexpr (VAR == VAR_COL || VAR == VAR_4A) {
        if (act) {
            cx<T>* dst = out + cg;
            const long long nd = a.out_nd;
#pragma unroll
            for (int e = 0; e < E; ++e) dst[(long long)(t + TT * e) * nd] = v[e];
        }
    } else {  // VAR_4B: re-map through LDS so lanes run along c on the store
        LdsRow<N, C::PADSH> li(c);
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) lds[li(t + TT * e)] = v[e];
        __syncthreads();
        const int c2 = tid % TPB, t2 = tid / TPB;
        LdsRow<N, C::PADSH> li2(c2);
        const long long cg2 = (long long)blockIdx.x * TPB + c2;
        if (cg2 < a.count) {
            cx<T>* dst = out + cg2;
            const long long nd = a.out_nd;
#pragma unroll
            for (int e = 0; e < E; ++e) dst[(long long)(t2 + TT * e) * nd] = lds[li2(t2 + TT * e)];
        }
    }
}

// ---- registry ----------------------------------------------------------------
struct KernelInfo {
    const void* fn;
    int prec;       // 0 = complex64, 1 = complex128
    int n;
    int var;
    int dir;        // -1 forward, +1 inverse
    int e;          // points per thread
    int tpb;        // transforms per block
    int threads;
    int lds_bytes;
    int nrad;
    int rad[8];
};

void register_kernel(const KernelInfo& k);

template <typename T, class S, int E, int TPB, int VAR, int DIR>
struct Registrar {
    Registrar() {
        using C = KCfg<T, S, E, TPB, VAR>;
        KernelInfo k{};
        k.fn = (const void*)&k_fft<T, S, E, TPB, VAR, DIR>;
        k.prec = sizeof(T) == 4 ? 0 : 1;
        k.n = S::N;
        k.var = VAR;
        k.dir = DIR;
        k.e = E;
        k.tpb = TPB;
        k.threads = C::THREADS;
        k.lds_bytes = C::LDS_ELEMS * (int)sizeof(cx<T>);
        k.nrad = S::P;
        for (int i = 0; i < S::P; ++i) k.rad[i] = S::rad[i];
        register_kernel(k);
    }
};

}  // namespace afft
