// Compile-time Stockham core: one transform of N = R0*R1*...*R(P-1) points is
// held by TT = N/E threads, E points per thread in registers.  Thread t owns
// points n = t + TT*e (e < E) before the first pass and after the last pass
// ("cyclic" layout), so the first loads and last stores of consecutive threads
// touch consecutive elements.  Between passes the points go through LDS.
//
// Pass p (radix R, Ns = R0*...*R(p-1)), butterflies j = t + TT*b (b < E/R):
//   in  : x[j + r*N/R], r < R             (register v[b + r*E/R])
//   tw  : x *= W_{Ns*R}^{(j mod Ns) * r}  (table, forward sign; conj for DIR=+1)
//   DFT_R in registers
//   out : y[(j/Ns)*Ns*R + (j mod Ns) + r*Ns]
// which is the self-sorting Stockham recursion: natural order in and out.
// The reference's radix-2 Stockham (internal/kernels/stockham.go:21-101)
// is the R=2 special case of the same recursion.
#pragma once
#include <utility>

#include "fft_common.hpp"

namespace afft {

template <int... Rs>
struct Sched {
    static constexpr int P = sizeof...(Rs);
    static constexpr int rad[P] = {Rs...};
    static constexpr int N = (1 * ... * Rs);
    static constexpr int ns(int p) {
        int s = 1;
        for (int i = 0; i < p; ++i) s *= rad[i];
        return s;
    }
    // offset (in entries) of pass p's twiddle block; pass 0 has none
    static constexpr int twoff(int p) {
        int o = 0;
        for (int i = 1; i < p; ++i) o += ns(i) * (rad[i] - 1);
        return o;
    }
    static constexpr int twsize() { return twoff(P); }
    static constexpr int maxrad() {
        int m = 0;
        for (int i = 0; i < P; ++i) m = rad[i] > m ? rad[i] : m;
        return m;
    }
};

// LDS addressing for transform-major ("row") blocks: transform c occupies a
// padded slab; one pad slot every 2^PADSH points spreads the strided writes of
// the early passes over all banks.
template <int N, int PADSH>
struct LdsRow {
    static constexpr int padded(int n) { return PADSH > 0 ? n + (n >> PADSH) : n; }
    static constexpr int SLAB = padded(N) + 1;   // one spare slot: point N for the real kernels
    int base;
    __device__ __forceinline__ LdsRow(int c) : base(c * SLAB) {}
    __device__ __forceinline__ int operator()(int n) const {
        return base + (PADSH > 0 ? n + (n >> PADSH) : n);
    }
};

// LDS addressing for column-interleaved blocks: point n of transform c at
// n*TPB + c, so lanes that differ in c hit consecutive banks.
template <int TPB>
struct LdsCol {
    int c;
    __device__ __forceinline__ LdsCol(int c_) : c(c_) {}
    __device__ __forceinline__ int operator()(int n) const { return n * TPB + c; }
};

template <typename T, class S, int E, int DIR, int p, bool HOIST, class TW, class L>
__device__ __forceinline__ void core_pass(cx<T> (&v)[E], cx<T>* lds, int t, TW tw, const L& li) {
    constexpr int N = S::N;
    constexpr int TT = N / E;
    constexpr int R = S::rad[p];
    constexpr int NS = S::ns(p);
    constexpr int B = E / R;
    static_assert(E % R == 0, "every radix must divide the points per thread");
    // r-major twiddle table: w^r of butterfly k at (r-1)*NS + k.  Radix 8
    // and 16 load w^1, w^2, w^4 (, w^8) and form the other powers as at most
    // three products: a quarter of the table traffic, at a cost of a few
    // ulps on the composite powers.  The loads depend only on the thread's
    // butterflies; with HOIST they are issued before the barrier that
    // precedes the pass's LDS reads, so their latency overlaps the wait.
    constexpr int NL = NS == 1 ? 1 : (R == 16 ? 4 : R == 8 ? 3 : R - 1);
    cx<T> wl[B][NL];
    auto load_tw = [&]() {
        if constexpr (NS > 1)
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const TW w = tw + S::twoff(p) + (t + TT * b) % NS;
            if constexpr (R == 16 || R == 8) {
                wl[b][0] = w[0];
                wl[b][1] = w[NS];
                wl[b][2] = w[3 * NS];
                if constexpr (R == 16) wl[b][3] = w[7 * NS];
            } else {
#pragma unroll
                for (int r = 1; r < R; ++r) wl[b][r - 1] = w[(r - 1) * NS];
            }
        }
    };
    if constexpr (NS > 1 && HOIST) load_tw();
    if constexpr (p > 0) {
        __syncthreads();
#pragma unroll
        for (int b = 0; b < B; ++b)
#pragma unroll
            for (int r = 0; r < R; ++r) v[b + r * B] = lds[li(t + TT * b + r * (N / R))];
    }
    if constexpr (NS > 1 && !HOIST) load_tw();
    if constexpr (NS > 1) {
#pragma unroll
        for (int b = 0; b < B; ++b) {
            if constexpr (R == 16 || R == 8) {
                cx<T> ww[R];
                ww[1] = wl[b][0];
                ww[2] = wl[b][1];
                ww[4] = wl[b][2];
                ww[3] = cmul(ww[1], ww[2]);
                ww[5] = cmul(ww[1], ww[4]);
                ww[6] = cmul(ww[2], ww[4]);
                ww[7] = cmul(ww[3], ww[4]);
                if constexpr (R == 16) {
                    ww[8] = wl[b][3];
#pragma unroll
                    for (int r = 9; r < 16; ++r) ww[r] = cmul(ww[r - 8], ww[8]);
                }
#pragma unroll
                for (int r = 1; r < R; ++r) v[b + r * B] = twmul<DIR>(v[b + r * B], ww[r]);
            } else {
#pragma unroll
                for (int r = 1; r < R; ++r) v[b + r * B] = twmul<DIR>(v[b + r * B], wl[b][r - 1]);
            }
        }
    }
#pragma unroll
    for (int b = 0; b < B; ++b) dft<R, DIR, B>(v + b);
    if constexpr (p < S::P - 1) {
        if constexpr (p > 0) __syncthreads();
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int j = t + TT * b;
            const int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
            for (int r = 0; r < R; ++r) lds[li(base + r * NS)] = v[b + r * B];
        }
    }
}

template <typename T, class S, int E, int DIR, bool HOIST, class TW, class L, int... Ps>
__device__ __forceinline__ void core_all(cx<T> (&v)[E], cx<T>* lds, int t, TW tw, const L& li,
                                         std::integer_sequence<int, Ps...>) {
    (core_pass<T, S, E, DIR, Ps, HOIST>(v, lds, t, tw, li), ...);
}

// Full transform.  PRESYNC adds a barrier before the first LDS write, needed
// when the LDS slab was read earlier in the same kernel.  HOIST issues each
// pass's twiddle loads before its barrier (see core_pass).  tw is the
// twiddle table: a const cx<T>* or a gptr<cx<T>>.
template <typename T, class S, int E, int DIR, bool PRESYNC = false, bool HOIST = false, class TW, class L>
__device__ __forceinline__ void core(cx<T> (&v)[E], cx<T>* lds, int t, TW tw, const L& li) {
    if constexpr (PRESYNC && S::P > 1) __syncthreads();
    core_all<T, S, E, DIR, HOIST>(v, lds, t, tw, li, std::make_integer_sequence<int, S::P>{});
}

}  // namespace afft
