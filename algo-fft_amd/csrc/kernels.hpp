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

enum { VAR_ROW = 0, VAR_COL = 1, VAR_4A = 2, VAR_4B = 3, VAR_COUNT = 4, VAR_BLUE = 4, VAR_REAL = 5 };

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
    int tw4_by_y;            // 1: the four-step column index is blockIdx.y, not c
    // Bluestein fusion on the four-step passes (runtime, uniform per launch):
    //   FUSE_CHIRP_IN (VAR_4A): x[n] * chirp[n] for n < flen, zero beyond
    //   FUSE_MUL_OUT  (VAR_4B): y[k] * fvec[k]
    //   FUSE_CROP_OUT (VAR_4B): y[k] * fvec[k] stored only for k < flen
    int fuse;
    int fconj;               // use conj(fvec)
    const void* fvec;
    long long flen;
    int nt;                  // NT_STORE / NT_LOAD: non-temporal output stores / input loads
};

enum { FUSE_NONE = 0, FUSE_CHIRP_IN = 1, FUSE_MUL_OUT = 2, FUSE_CROP_OUT = 3 };

template <typename T>
__device__ __forceinline__ cx<T> fvec_at(const void* f, long long i, int conj) {
    cx<T> w = ((const cx<T>*)f)[i];
    return conj ? conjx(w) : w;
}

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

    if constexpr (VAR == VAR_ROW) {
        const cx<T>* src = in + cg * a.in_cd;
        if (a.nt & NT_LOAD) {
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = act ? ld_nt(src + t + TT * e) : mk<T>(0, 0);
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = act ? src[t + TT * e] : mk<T>(0, 0);
        }
    } else if constexpr (VAR == VAR_4B) {
        const cx<T>* src = in + cg * a.in_cd;
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = act ? src[t + TT * e] : mk<T>(0, 0);
    } else if (VAR == VAR_4A && a.fuse == FUSE_CHIRP_IN) {
        // element n of the padded sequence sits at offset n of a unit-stride row
        const long long nd = a.in_nd;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const long long n = cg + (long long)(t + TT * e) * nd;
            v[e] = (act && n < a.flen) ? cmul(in[n], fvec_at<T>(a.fvec, n, a.fconj)) : mk<T>(0, 0);
        }
    } else {
        const cx<T>* src = in + cg;
        const long long nd = a.in_nd;
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = act ? src[(long long)(t + TT * e) * nd] : mk<T>(0, 0);
    }

    if constexpr (C::ROWL) {
        LdsRow<N, C::PADSH> li(c);
        // one row per block of at most 256 threads: twiddle loads ahead of
        // each pass's barrier.  Measured on 512 MiB batches: complex64
        // N=4096 0.182 -> 0.177 ms, complex128 N=4096 0.189 -> 0.187 ms;
        // with several rows per block (N=256), in the four-step passes and
        // in the 512- and 1024-thread rows (complex128 8192, complex64
        // 16384) it lost 1-2 %.
        core<T, S, E, DIR, false, VAR == VAR_ROW && TPB == 1 && N / E <= 256>(v, lds, t, tw, li);
    } else {
        LdsCol<TPB> li(c);
        core<T, S, E, DIR>(v, lds, t, tw, li);
    }

    if constexpr (VAR == VAR_4A) {
        const long long col = (a.tw4_by_y ? (long long)blockIdx.y : cg) + a.tw4_c0;
        // W_L^{k*col} for k = t + TT*e (k < N and col < L/N, so k*col < L
        // needs no reduction).  Lanes differ in col, so every table load
        // gathers from scattered lines: look up every fourth e and step
        // between them by W_L^{TT*col}, at most three products from a
        // table value (10 table loads per thread instead of 2E).
        if constexpr (E % 4 == 0) {
            const cx<T> st = tw4<T>(a, (long long)TT * col);
#pragma unroll
            for (int e0 = 0; e0 < E; e0 += 4) {
                cx<T> w = tw4<T>(a, (long long)(t + TT * e0) * col);
                v[e0] = twmul<DIR>(v[e0], w);
#pragma unroll
                for (int e = e0 + 1; e < e0 + 4; ++e) {
                    w = cmul(w, st);
                    v[e] = twmul<DIR>(v[e], w);
                }
            }
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = twmul<DIR>(v[e], tw4<T>(a, (long long)(t + TT * e) * col));
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
            for (int e = 0; e < E; ++e) st_out(dst + t + TT * e, v[e], a.nt);
        }
    } else if constexpr (VAR == VAR_COL || VAR == VAR_4A) {
        if (act) {
            cx<T>* dst = out + cg;
            const long long nd = a.out_nd;
#pragma unroll
            for (int e = 0; e < E; ++e) st_out(dst + (long long)(t + TT * e) * nd, v[e], a.nt);
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
            const long long nd = a.out_nd;
            if (a.fuse == FUSE_NONE) {
#pragma unroll
                for (int e = 0; e < E; ++e) st_out(out + cg2 + (long long)(t2 + TT * e) * nd, lds[li2(t2 + TT * e)], a.nt);
            } else {
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const long long k = cg2 + (long long)(t2 + TT * e) * nd;
                    const cx<T> y = cmul(lds[li2(t2 + TT * e)], fvec_at<T>(a.fvec, k < a.flen ? k : 0, a.fconj));
                    if (a.fuse == FUSE_MUL_OUT || k < a.flen) out[k] = y;
                }
            }
        }
    }
}

// ---- fused Bluestein chain for a padded length M held in LDS ---------------
// a[n] = x[n]*chirp[n] (n < N, zero to M), A = FFT_M(a), A *= F (the filter
// spectrum with 1/M folded in), y = IFFT_M(A), X[k] = y[k]*chirp[k]*scale.
// The inverse transform uses conj(chirp) and its own filter.  Both FFTs run
// on the register/LDS core; one global read and one global write per point
// (plan_bluestein.go:7-58, internal/kernels/bluestein.go:39-77).
struct BArgs {
    const void* in;
    void* out;
    const void* tw;       // schedule twiddles of M
    const void* chirp;    // N entries, forward sign
    const void* filt;     // M entries
    long long count;
    long long in_cd, out_cd;
    long long nlog;
    double scale;
};

template <typename T, class S, int E, int TPB, int DIR>
__global__ void __launch_bounds__((KCfg<T, S, E, TPB, VAR_ROW>::THREADS)) k_blue(BArgs a) {
    using C = KCfg<T, S, E, TPB, VAR_ROW>;
    constexpr int TT = C::TT;
    __shared__ cx<T> lds[C::LDS_ELEMS];
    const int tid = threadIdx.x;
    const int c = tid / TT, t = tid % TT;
    const long long cg = (long long)blockIdx.x * TPB + c;
    const bool act = cg < a.count;
    const cx<T>* __restrict__ src = (const cx<T>*)a.in + cg * a.in_cd;
    const cx<T>* __restrict__ chirp = (const cx<T>*)a.chirp;
    const cx<T>* __restrict__ filt = (const cx<T>*)a.filt;
    const cx<T>* __restrict__ tw = (const cx<T>*)a.tw;
    cx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int n = t + TT * e;
        v[e] = (act && n < a.nlog) ? twmul<DIR>(src[n], chirp[n]) : mk<T>(0, 0);
    }
    LdsRow<S::N, C::PADSH> li(c);
    core<T, S, E, -1>(v, lds, t, tw, li);
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = cmul(v[e], filt[t + TT * e]);
    core<T, S, E, 1, true>(v, lds, t, tw, li);
    if (act) {
        cx<T>* dst = (cx<T>*)a.out + cg * a.out_cd;
        const T sc = (T)a.scale;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int n = t + TT * e;
            if (n < a.nlog) dst[n] = scl(twmul<DIR>(v[e], chirp[n]), sc);
        }
    }
}

// ---- real transforms: pack/unpack fused into the first/last pass --------------
// R2C (PlanRealT.forwardSingle, plan_real_generic.go:142-228): the N reals of
// a row are read as h = N/2 complex points z[m] = x[2m] + i x[2m+1]; after the
// h-point FFT Z sits in LDS and each thread recombines its bins
//   X[k] = ½(A + B) - ½ i W_N^k (A - B),   A = Z[k], B = conj(Z[h-k])
// (the reference's X = A - U[k](A - B) with U[k] = ½(1 + i W_N^k)), with
// X[0] = Re Z0 + Im Z0 and X[h] = Re Z0 - Im Z0.
// C2R (inverseSingle :289-355 with RepackInverse, real_repack.go:49-75):
// the h+1 bins go through LDS, each thread forms
//   Z[k] = ½(A + B) + ½ i conj(W_N^k) (A - B),   A = X[k], B = conj(X[h-k])
// (imaginary parts of X[0], X[h] ignored, as the reference uses only their
// real parts), then the inverse h-point FFT scaled by 1/h gives x as pairs.
struct RArgs {
    const void* in;
    void* out;
    const void* tw;       // schedule twiddles of h
    const void* wk;       // W_N^k, k < h (forward sign)
    long long count;
    long long in_d, out_d;   // per-row distance in complex elements of each side
    double scale;
};

template <typename T, class S, int E, int TPB, int DIR>
__global__ void __launch_bounds__((KCfg<T, S, E, TPB, VAR_ROW>::THREADS)) k_real(RArgs a) {
    using C = KCfg<T, S, E, TPB, VAR_ROW>;
    constexpr int H = S::N;
    constexpr int TT = C::TT;
    using L = LdsRow<H, C::PADSH>;
    __shared__ cx<T> lds[TPB * L::SLAB];
    const int tid = threadIdx.x;
    const int c = tid / TT, t = tid % TT;
    const long long cg = (long long)blockIdx.x * TPB + c;
    const bool act = cg < a.count;
    const cx<T>* __restrict__ src = (const cx<T>*)a.in + cg * a.in_d;
    cx<T>* __restrict__ dst = (cx<T>*)a.out + cg * a.out_d;
    const cx<T>* __restrict__ wk = (const cx<T>*)a.wk;
    const T half = T(0.5);
    const T sc = (T)a.scale;
    L li(c);
    cx<T> v[E];
    if constexpr (DIR < 0) {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = act ? src[t + TT * e] : mk<T>(0, 0);
        core<T, S, E, -1>(v, lds, t, (const cx<T>*)a.tw, li);
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) lds[li(t + TT * e)] = v[e];
        __syncthreads();
        if (act) {
            // W_N^k for k = t + TT*e: a table value every fourth e and the
            // step W_N^TT in between (at most three products).  Measured on
            // float32 R2C 4096 x 32768: 0.230 -> 0.200 ms against a table
            // load per bin; C2R 512^3 0.788 -> 0.765 ms.
            const cx<T> st = wk[TT];
            cx<T> w = mk<T>(1, 0);
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int k = t + TT * e;
                w = e % 4 == 0 ? wk[k] : cmul(w, st);
                const cx<T> A = lds[li(k)];
                if (k == 0) {
                    dst[0] = mk<T>((A.x + A.y) * sc, T(0));
                    dst[H] = mk<T>((A.x - A.y) * sc, T(0));
                } else {
                    const cx<T> B = conjx(lds[li(H - k)]);
                    const cx<T> S2 = scl(A + B, half), D = scl(A - B, half);
                    const cx<T> X = S2 + rot<-1>(cmul(w, D));
                    dst[k] = scl(X, sc);
                }
            }
        }
    } else {
        for (int k = t; k <= H; k += TT) lds[li(k)] = act ? src[k] : mk<T>(0, 0);
        __syncthreads();
        const cx<T> st = wk[TT];
        cx<T> w = mk<T>(1, 0);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int k = t + TT * e;
            w = e % 4 == 0 ? wk[k] : cmul(w, st);
            cx<T> A = lds[li(k)];
            cx<T> B = conjx(lds[li(H - k)]);
            if (k == 0) { A.y = T(0); B.y = T(0); }
            const cx<T> S2 = scl(A + B, half), D = scl(A - B, half);
            v[e] = S2 + rot<1>(cmulc(D, w));   // + i conj(W) D
        }
        core<T, S, E, 1, true>(v, lds, t, (const cx<T>*)a.tw, li);
        if (act) {
#pragma unroll
            for (int e = 0; e < E; ++e) dst[t + TT * e] = scl(v[e], sc);
        }
    }
}

// Stand-alone recombination for the sizes whose h-point FFT runs as its own
// plan step: in = Z (h complex per row), out = X (h+1 per row) for R2C;
// in = X, out = Z for C2R.  One thread per bin.
struct RPArgs {
    const void* in;
    void* out;
    const void* wk;
    long long rows, h;
    long long in_d, out_d;
    double scale;
};
template <typename T, int DIR>
__global__ void __launch_bounds__(256) k_rpost(RPArgs a) {
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long per = DIR < 0 ? a.h + 1 : a.h;
    if (idx >= a.rows * per) return;
    const long long r = idx / per, k = idx - r * per;
    const cx<T>* in = (const cx<T>*)a.in + r * a.in_d;
    cx<T>* out = (cx<T>*)a.out + r * a.out_d;
    const cx<T>* wk = (const cx<T>*)a.wk;
    const T half = T(0.5), sc = (T)a.scale;
    if constexpr (DIR < 0) {
        if (k == 0 || k == a.h) {
            const cx<T> z = in[0];
            out[k] = mk<T>((k == 0 ? z.x + z.y : z.x - z.y) * sc, T(0));
            return;
        }
        const cx<T> A = in[k], B = conjx(in[a.h - k]);
        const cx<T> S2 = scl(A + B, half), D = scl(A - B, half);
        out[k] = scl(S2 + rot<-1>(cmul(wk[k], D)), sc);
    } else {
        cx<T> A = in[k], B = conjx(in[a.h - k]);
        if (k == 0) { A.y = T(0); B.y = T(0); }
        const cx<T> S2 = scl(A + B, half), D = scl(A - B, half);
        out[k] = scl(S2 + rot<1>(cmulc(D, wk[k])), sc);
    }
}

// Strided 2-D copy: element i of line c from in[c*in_cd + i*in_nd] to
// out[c*out_cd + i*out_nd].  Used for length-1 axes and to gather strided
// lines into contiguous rows (and scatter them back) for the paths that need
// unit stride.
struct CopyArgs {
    const void* in;
    void* out;
    long long lines, len;
    long long in_cd, in_nd, out_cd, out_nd;
    double scale;
};
template <typename T>
__global__ void __launch_bounds__(256) k_copy(CopyArgs a) {
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= a.lines * a.len) return;
    const long long c = idx / a.len, i = idx - c * a.len;
    cx<T> v = ((const cx<T>*)a.in)[c * a.in_cd + i * a.in_nd];
    if (a.scale != 1.0) v = scl(v, (T)a.scale);
    ((cx<T>*)a.out)[c * a.out_cd + i * a.out_nd] = v;
}

// ---- registry ----------------------------------------------------------------
enum { KIND_FFT = 0, KIND_BLUE = 1, KIND_REAL = 2 };

struct KernelInfo {
    const void* fn;
    int kind;       // KIND_FFT (k_fft) or KIND_BLUE (k_blue)
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
        using C = KCfg<T, S, E, TPB, (VAR >= VAR_BLUE ? VAR_ROW : VAR)>;
        KernelInfo k{};
        if constexpr (VAR == VAR_BLUE) {
            k.fn = (const void*)&k_blue<T, S, E, TPB, DIR>;
            k.kind = KIND_BLUE;
        } else if constexpr (VAR == VAR_REAL) {
            k.fn = (const void*)&k_real<T, S, E, TPB, DIR>;
            k.kind = KIND_REAL;
        } else {
            k.fn = (const void*)&k_fft<T, S, E, TPB, VAR, DIR>;
            k.kind = KIND_FFT;
        }
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
