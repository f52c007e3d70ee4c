// Runtime-size Stockham kernel for 2/3/5/7-smooth lengths (the mixed-radix
// path; the reference's mixedradix.go:45-358 does the same recursion
// depth-first on the CPU and sends 7-smooth sizes to Bluestein instead).
//
// A block of 256 threads holds G = blockDim.x / TPT transforms, TPT threads
// each: about 8 points per thread for rows, 16 for columns.  Rows above 4096
// points take one transform per block of 1024 threads (complex64) or 512
// (complex128).  Every pass reads its butterfly inputs from one LDS buffer
// (global memory for the first pass), applies the per-pass twiddles, runs
// the radix-R codelet and writes to the other LDS buffer (global memory for
// the last pass, whose output index j + r*Ns is already natural order).  One
// barrier per pass.
#pragma once
#include "kernels.hpp"

namespace afft {

constexpr int GEN_THREADS = 256;      // threads per block up to 8 * 256 points per transform
constexpr int GEN_MAXTHREADS = 512;   // threads per block above
constexpr int GEN_ROWTHREADS = 1024;  // complex64 rows above 4096 points: one transform, up to 8 points a thread
constexpr int GEN_MAXPASS = 16;

struct GArgs {
    KArgs k;
    int n;
    int tpt;                     // threads per transform (power of two)
    int npass;
    int rad[GEN_MAXPASS];
    int ns[GEN_MAXPASS];
    unsigned nsm[GEN_MAXPASS];   // mod_magic(ns[p])
    int twoff[GEN_MAXPASS];
};

template <int DIR, typename T>
__device__ __forceinline__ void dft_rt(int R, cx<T>* v) {
    switch (R) {
    case 2: dft<2, DIR, 1>(v); break;
    case 3: dft<3, DIR, 1>(v); break;
    case 4: dft<4, DIR, 1>(v); break;
    case 5: dft<5, DIR, 1>(v); break;
    case 7: dft<7, DIR, 1>(v); break;
    case 8: dft<8, DIR, 1>(v); break;
    case 16: dft<16, DIR, 1>(v); break;
    default: break;
    }
}

// Twiddles w^1 .. w^(R-1) of one butterfly from its k-major table entries
// w[r-1] = w^r.  For complex128, radices 4 and up load w^1, w^2 (and w^4,
// w^8 for radix 8 and 16) and form the other powers as at most two products
// (a few ulps): the table loads of a pass drop from one per point to about
// one per four points for radix 16.  Measured on 512 MiB batches:
// complex128 N=1000 0.287 -> 0.278 ms, N=2500 0.292 -> 0.277 ms, N=3000
// 0.326 -> 0.302 ms, N=6000 0.408 -> 0.377 ms, N=100000 0.981 -> 0.957 ms.  Complex64 passes (tw_apply) load every power as they
// apply it: the products' registers pushed the in-place rows from 23 to
// 39-43 spilled VGPRs and the two-buffer columns from 3 to 2 waves per SIMD
// (N=6000 0.365 -> 0.436 ms, the ragged set 0.499 -> 0.562 ms).
template <int R, typename T, class W>
__device__ __forceinline__ void tw_powers(W w, cx<T> (&ww)[R]) {
    if constexpr (R <= 3) {
#pragma unroll
        for (int r = 1; r < R; ++r) ww[r] = w[r - 1];
    } else if constexpr (R == 4 || R == 5 || R == 7) {
        ww[1] = w[0];
        ww[2] = w[1];
        ww[3] = cmul(ww[1], ww[2]);
        if constexpr (R >= 5) ww[4] = cmul(ww[2], ww[2]);
        if constexpr (R == 7) {
            ww[5] = cmul(ww[4], ww[1]);
            ww[6] = cmul(ww[4], ww[2]);
        }
    } else {
        static_assert(R == 8 || R == 16, "radix");
        ww[1] = w[0];
        ww[2] = w[1];
        ww[4] = w[3];
        ww[3] = cmul(ww[1], ww[2]);
        ww[5] = cmul(ww[1], ww[4]);
        ww[6] = cmul(ww[2], ww[4]);
        ww[7] = cmul(ww[3], ww[4]);
        if constexpr (R == 16) {
            ww[8] = w[7];
#pragma unroll
            for (int r = 9; r < 16; ++r) ww[r] = cmul(ww[r - 8], ww[8]);
        }
    }
}

// v[r] *= w^r (r = 1 .. R-1) for the butterfly whose table entries start at w.
template <int R, int DIR, typename T, class W>
__device__ __forceinline__ void tw_apply(W w, cx<T>* v) {
    if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = twmul<DIR>(v[r], w[r - 1]);
    } else {
        cx<T> ww[R];
        tw_powers<R>(w, ww);
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = twmul<DIR>(v[r], ww[r]);
    }
}

// One runtime Stockham pass with a compile-time radix R, so the butterfly
// stays in registers: thread t of T handles butterflies j = t, t+T, ... of
// the N/R in the pass; load(n) gives input point n, store(k, v, r) takes
// output point k = (j/Ns)*Ns*R + (j mod Ns) + r*Ns (r counts 0 .. R-1 within
// the butterfly).
template <int R, int DIR, typename T, class W, class Load, class Store>
__device__ __forceinline__ void rt_pass(int t, int TT, int N, int NS, unsigned NSM, W tw, Load load, Store store) {
    const int NR = N / R;
    for (int j = t; j < NR; j += TT) {
        cx<T> v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = load(j + r * NR);
        const int k = fast_mod(j, NS, NSM);
        if (NS > 1) tw_apply<R, DIR>(tw + k * (R - 1), v);   // k-major table: k*(R-1) + r-1
        dft<R, DIR, 1>(v);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) store(base + r * NS, v[r], r);
    }
}

template <int DIR, typename T, class W, class Load, class Store>
__device__ __forceinline__ void rt_pass_any(int R, int t, int TT, int N, int NS, unsigned NSM, W tw, Load load,
                                            Store store) {
    switch (R) {
    case 2: rt_pass<2, DIR, T>(t, TT, N, NS, NSM, tw, load, store); break;
    case 3: rt_pass<3, DIR, T>(t, TT, N, NS, NSM, tw, load, store); break;
    case 4: rt_pass<4, DIR, T>(t, TT, N, NS, NSM, tw, load, store); break;
    case 5: rt_pass<5, DIR, T>(t, TT, N, NS, NSM, tw, load, store); break;
    case 7: rt_pass<7, DIR, T>(t, TT, N, NS, NSM, tw, load, store); break;
    case 8: rt_pass<8, DIR, T>(t, TT, N, NS, NSM, tw, load, store); break;
    case 16: rt_pass<16, DIR, T>(t, TT, N, NS, NSM, tw, load, store); break;
    default: break;
    }
}

// One in-place radix-R Stockham pass of a transform of N points with
// N/R <= (16/R) * TT: thread t reads butterflies j = t, t+TT, ... (load(n)),
// applies twiddles and the codelet, and after a barrier of the whole block
// writes output point (j/Ns)*Ns*R + j%Ns + r*Ns (store(k, v)).  Every thread
// of the block calls this; `active` threads do the work.
template <int R, int DIR, typename T, class W, class Load, class Store>
__device__ __forceinline__ void rt_pass_ip(bool active, int t, int TT, int N, int NS, unsigned NSM, W tw, Load load,
                                           Store store) {
    constexpr int MB = (16 + R - 1) / R;
    const int NR = N / R;
    cx<T> v[MB][R];
    if (active) {
#pragma unroll
        for (int b = 0; b < MB; ++b) {
            const int j = t + b * TT;
            if (j < NR) {
#pragma unroll
                for (int r = 0; r < R; ++r) v[b][r] = load(j + r * NR);
                const int k = fast_mod(j, NS, NSM);
                if (NS > 1) tw_apply<R, DIR>(tw + k * (R - 1), v[b]);   // k-major table: k*(R-1) + r-1
                dft<R, DIR, 1>(v[b]);
            }
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int b = 0; b < MB; ++b) {
            const int j = t + b * TT;
            if (j < NR) {
                const int k = fast_mod(j, NS, NSM);
                const int base = (j - k) * R + k;
#pragma unroll
                for (int r = 0; r < R; ++r) store(base + r * NS, v[b][r]);
            }
        }
    }
}

template <int DIR, typename T, class W, class Load, class Store>
__device__ __forceinline__ void rt_pass_ip_any(int R, bool active, int t, int TT, int N, int NS, unsigned NSM, W tw,
                                               Load load, Store store) {
    switch (R) {
    case 2: rt_pass_ip<2, DIR, T>(active, t, TT, N, NS, NSM, tw, load, store); break;
    case 3: rt_pass_ip<3, DIR, T>(active, t, TT, N, NS, NSM, tw, load, store); break;
    case 4: rt_pass_ip<4, DIR, T>(active, t, TT, N, NS, NSM, tw, load, store); break;
    case 5: rt_pass_ip<5, DIR, T>(active, t, TT, N, NS, NSM, tw, load, store); break;
    case 7: rt_pass_ip<7, DIR, T>(active, t, TT, N, NS, NSM, tw, load, store); break;
    case 8: rt_pass_ip<8, DIR, T>(active, t, TT, N, NS, NSM, tw, load, store); break;
    case 16: rt_pass_ip<16, DIR, T>(active, t, TT, N, NS, NSM, tw, load, store); break;
    default: {   // radix 1: the copy of a one-point transform
        cx<T> v0 = mk<T>(0, 0);
        if (active && t < N) v0 = load(t);
        __syncthreads();
        if (active && t < N) store(t, v0);
    }
    }
}

// VAR_ROW / VAR_COL / VAR_4A / VAR_4B as in kernels.hpp.  ROW-like variants
// lay a transform out contiguously in LDS, COL variants interleave G columns.
template <typename T, int VAR, int DIR>
__global__ void __launch_bounds__((VAR == VAR_ROW && sizeof(T) == 4 ? GEN_ROWTHREADS : GEN_MAXTHREADS)) k_gen(GArgs g) {
    extern __shared__ unsigned char smraw[];
    cx<T>* sm = (cx<T>*)smraw;
    const KArgs a = g.k;
    const int N = g.n;
    const int TPT = g.tpt;
    const int lgt = __ffs(TPT) - 1;              // TPT and G are powers of two
    const int G = blockDim.x >> lgt;
    const int tid = threadIdx.x;
    constexpr bool ROWL = (VAR == VAR_ROW || VAR == VAR_4B);
    int c, t;
    if constexpr (ROWL) { c = tid >> lgt; t = tid & (TPT - 1); }
    else { c = tid & (G - 1); t = tid >> (__ffs(G) - 1); }
    const long long cg = (long long)blockIdx.x * G + c;
    const bool act = cg < a.count;
    auto lidx = [&](int buf, int n) -> int {
        if constexpr (ROWL) return buf * G * N + c * N + n;
        else return buf * G * N + n * G + c;
    };
    const cx<T>* __restrict__ in = (const cx<T>*)a.in + (long long)blockIdx.y * a.in_bd;
    cx<T>* __restrict__ out = (cx<T>*)a.out + (long long)blockIdx.y * a.out_bd;
    const cx<T>* __restrict__ tw = (const cx<T>*)a.tw;
    const cx<T>* src_g = (VAR == VAR_ROW || VAR == VAR_4B) ? in + cg * a.in_cd : in + cg;
    const long long ind = (VAR == VAR_ROW || VAR == VAR_4B) ? 1 : a.in_nd;

    for (int p = 0; p < g.npass; ++p) {
        const int R = g.rad[p];
        const int NS = g.ns[p];
        const unsigned NSM = g.nsm[p];
        const bool last = p == g.npass - 1;
        const cx<T>* twp = tw + g.twoff[p];
        const int rb = (p - 1) & 1, wb = p & 1;
        // separate lambdas (and rt_pass instances) for global and LDS sides,
        // so no store or load has to choose its address space at run time
        auto load_g = [&](int n) -> cx<T> { return act ? src_g[(long long)n * ind] : mk<T>(0, 0); };
        auto load_s = [&](int n) -> cx<T> { return sm[lidx(rb, n)]; };
        auto store_s = [&](int k, cx<T> y, int) { sm[lidx(wb, k)] = y; };
        // VAR_4A: W_L^{k*col} for the outputs k = base + r*Ns of a butterfly:
        // a table value every fourth r, the step W_L^{Ns*col} in between
        // (lanes differ in col, so the table gathers hit scattered lines)
        const long long col = cg + a.tw4_c0;
        cx<T> tw_step = mk<T>(1, 0), tw_cur = mk<T>(1, 0);
        if constexpr (VAR == VAR_4A)
            if (last) tw_step = tw4<T>(a, (long long)NS * col);
        auto store_o = [&](int k, cx<T> y, int r) {
            if constexpr (VAR == VAR_4A) {
                tw_cur = (r % 4 == 0) ? tw4<T>(a, (long long)k * col) : cmul(tw_cur, tw_step);
                y = twmul<DIR>(y, tw_cur);
            }
            if (a.scale != 1.0) y = scl(y, (T)a.scale);
            if constexpr (VAR == VAR_ROW) {
                if (act) st_out(out + cg * a.out_cd + k, y, a.nt);
            } else if constexpr (VAR == VAR_COL || VAR == VAR_4A) {
                if (act) st_out(out + cg + (long long)k * a.out_nd, y, a.nt);
            } else {
                sm[lidx(wb, k)] = y;   // re-mapped below
            }
        };
        if (p == 0) {
            if (last) rt_pass_any<DIR, T>(R, t, TPT, N, NS, NSM, twp, load_g, store_o);
            else rt_pass_any<DIR, T>(R, t, TPT, N, NS, NSM, twp, load_g, store_s);
        } else {
            if (last) rt_pass_any<DIR, T>(R, t, TPT, N, NS, NSM, twp, load_s, store_o);
            else rt_pass_any<DIR, T>(R, t, TPT, N, NS, NSM, twp, load_s, store_s);
        }
        __syncthreads();
    }
    if constexpr (VAR == VAR_4B) {
        // transposing store: lanes run along the transform index c
        const int fb = (g.npass - 1) & 1;
        const int c2 = tid & (G - 1), t2 = tid >> (__ffs(G) - 1);
        const long long cg2 = (long long)blockIdx.x * G + c2;
        if (cg2 < a.count) {
            for (int k = t2; k < N; k += TPT) st_out(out + cg2 + (long long)k * a.out_nd, sm[fb * G * N + c2 * N + k], a.nt);
        }
    }
}

// Rows of one transform per block, every pass in place in ONE LDS buffer of
// N points: thread t of TT = blockDim.x (a multiple of 64, TT >= N/16) reads
// all its butterflies of a pass, the block passes a barrier, then it writes.
// Half the LDS of the two-buffer form: complex64 rows up to 16384 points
// (128 KiB) and complex128 up to 8192 fit one pass, and shorter rows leave
// room for more blocks per CU.  The input is staged into LDS first (every
// load of a thread in flight at once): loading it inside the first pass, as
// the two-buffer kernel does, took the kernel to 128 VGPRs plus 325 spilled
// at 1024 threads per block.  The last pass stores to global memory.
template <typename T>
struct GenIP {
    static constexpr int THREADS = sizeof(T) == 4 ? 1024 : 512;
    // waves per SIMD at the register use of k_gen_ip (128 VGPRs for
    // complex64 at 1024 threads, about 200 for complex128 at 512)
    static constexpr int WAVES_PER_SIMD = sizeof(T) == 4 ? 4 : 2;
};

template <typename T, int DIR>
__global__ void __launch_bounds__(GenIP<T>::THREADS) k_gen_ip(GArgs g) {
    extern __shared__ unsigned char smraw[];
    cx<T>* sm = (cx<T>*)smraw;
    const KArgs a = g.k;
    const int N = g.n;
    const int TT = blockDim.x;
    const int t = threadIdx.x;
    const long long cg = blockIdx.x;
    const cx<T>* __restrict__ src = (const cx<T>*)a.in + (long long)blockIdx.y * a.in_bd + cg * a.in_cd;
    cx<T>* __restrict__ dst = (cx<T>*)a.out + (long long)blockIdx.y * a.out_bd + cg * a.out_cd;
    const cx<T>* __restrict__ tw = (const cx<T>*)a.tw;
    {
        const bool ntl = a.nt & NT_LOAD;
        cx<T> v[16];
#pragma unroll
        for (int h = 0; h < 16; ++h) {
            const int n = t + h * TT;
            v[h] = n < N ? (ntl ? ld_nt(src + n) : src[n]) : mk<T>(0, 0);
        }
#pragma unroll
        for (int h = 0; h < 16; ++h)
            if (t + h * TT < N) sm[t + h * TT] = v[h];
    }
    __syncthreads();
    for (int p = 0; p < g.npass; ++p) {
        const int R = g.rad[p];
        const int NS = g.ns[p];
        const unsigned NSM = g.nsm[p];
        const cx<T>* twp = tw + g.twoff[p];
        auto load_s = [&](int n) -> cx<T> { return sm[n]; };
        auto store_s = [&](int k, cx<T> y) { sm[k] = y; };
        auto store_o = [&](int k, cx<T> y) {
            if (a.scale != 1.0) y = scl(y, (T)a.scale);
            st_out(dst + k, y, a.nt);
        };
        if (p == g.npass - 1) {
            rt_pass_ip_any<DIR, T>(R, true, t, TT, N, NS, NSM, twp, load_s, store_o);
        } else {
            rt_pass_ip_any<DIR, T>(R, true, t, TT, N, NS, NSM, twp, load_s, store_s);
            __syncthreads();
        }
    }
}

// Column layouts (VAR_COL strided axes, the four-step passes VAR_4A with
// the four-step twiddle and VAR_4B rows with the transposing store) with
// G = blockDim.x / TPT transforms per block and
// every pass in place in ONE LDS buffer of G*N points, for N <= 16*TPT: each
// thread holds all its butterflies of a pass in registers across the pass's
// barrier (rt_pass_ip).  Half the LDS of k_gen lets twice the blocks share
// a CU: k_gen's column passes hold 2-3 blocks of 4 waves per CU, latency
// bound with the vector ALUs under half busy.  The input is staged into LDS
// first with up to 16 loads of a thread in flight.
template <typename T>
struct GenIPG {
    // largest block: blocks grow to 512 threads where 256 would hold fewer
    // than 16 transforms, for longer column segments
    static constexpr int THREADS = 2 * GEN_THREADS;
    // waves per SIMD the register budget allows (128 VGPRs for complex64,
    // 256 for complex128)
    static constexpr int WAVES_PER_SIMD = sizeof(T) == 4 ? 4 : 2;
};

template <typename T, int VAR, int DIR>
__global__ void __launch_bounds__(GenIPG<T>::THREADS, GenIPG<T>::WAVES_PER_SIMD) k_gen_ipg(GArgs g) {
    extern __shared__ unsigned char smraw[];
    cx<T>* sm = (cx<T>*)smraw;
    const KArgs a = g.k;
    const int N = g.n;
    const int TPT = g.tpt;
    const int lgt = __ffs(TPT) - 1;
    const int lgg = __ffs(blockDim.x) - 1 - lgt;   // blockDim.x is a power of two
    const int G = 1 << lgg;
    static_assert(VAR == VAR_4A || VAR == VAR_4B || VAR == VAR_COL, "column layouts only");
    const int tid = threadIdx.x;
    // Every layout holds transform c at sm[c*N ..] (column-interleaved
    // slots made the column pass's addressing spill).  The passes run TPT
    // consecutive threads per transform; loads and stores of strided columns
    // (VAR_COL and VAR_4A input, every output) run lanes along c, so that
    // neighbouring lanes touch neighbouring addresses.
    const int c = tid >> lgt, t = tid & (TPT - 1);
    const int c2 = tid & (G - 1), t2 = tid >> lgg;
    // Column reads (VAR_COL, VAR_4A): consecutive blocks go to different
    // XCDs (each with its own L2) in turn, so neighbouring column groups,
    // which share cache lines, would be fetched by different L2s.  A block
    // takes the work item of its position within its XCD's share instead
    // (a bijection of the grid), so an XCD's blocks hold neighbouring
    // groups.  Measured on 512 MiB batches: complex64 1200 x 1500 0.811 ->
    // 0.673 ms, N=250000 0.808 -> 0.754 ms, complex128 1200 x 1500 0.657
    // -> 0.630 ms; complex64 1000 x 1000 0.569 -> 0.581 ms and N=24000
    // 0.563 -> 0.577 ms lost.
    long long bx = blockIdx.x, by = blockIdx.y;
    if constexpr (VAR != VAR_4B) {
        constexpr long long XCDS = 8;
        const long long tot = (long long)gridDim.x * gridDim.y, lin = by * gridDim.x + bx;
        const long long q = tot / XCDS, r = tot % XCDS, x = lin % XCDS;
        const long long nl = x * q + (x < r ? x : r) + lin / XCDS;
        bx = nl % gridDim.x;
        by = nl / gridDim.x;
    }
    const long long b0 = bx * G;
    const cx<T>* __restrict__ in = (const cx<T>*)a.in + by * a.in_bd;
    cx<T>* __restrict__ out = (cx<T>*)a.out + by * a.out_bd;
    const cx<T>* __restrict__ tw = (const cx<T>*)a.tw;
    constexpr bool ROWIN = VAR == VAR_4B;
    {
        const int sc = ROWIN ? c : c2, st = ROWIN ? t : t2;
        const bool act = b0 + sc < a.count;
        const bool ntl = a.nt & NT_LOAD;
        const cx<T>* q = ROWIN ? in + (b0 + sc) * a.in_cd + st : in + b0 + sc + (long long)st * a.in_nd;
        const long long step = ROWIN ? (long long)TPT : (long long)TPT * a.in_nd;
        cx<T> v[16];
#pragma unroll
        for (int h = 0; h < 16; ++h, q += step) {
            const int n = st + h * TPT;
            v[h] = (act && n < N) ? (ntl ? ld_nt(q) : *q) : mk<T>(0, 0);
        }
#pragma unroll
        for (int h = 0; h < 16; ++h)
            if (st + h * TPT < N) sm[sc * N + st + h * TPT] = v[h];
    }
    __syncthreads();
    const T scale = (T)a.scale;
    cx<T>* row = sm + c * N;
    for (int p = 0; p < g.npass; ++p) {
        const int R = g.rad[p];
        const int NS = g.ns[p];
        const unsigned NSM = g.nsm[p];
        const cx<T>* twp = tw + g.twoff[p];
        auto load_s = [&](int n) -> cx<T> { return row[n]; };
        auto store_s = [&](int k, cx<T> y) { row[k] = y; };
        rt_pass_ip_any<DIR, T>(R, true, t, TPT, N, NS, NSM, twp, load_s, store_s);
        if (p < g.npass - 1) __syncthreads();
    }
    {
        // column stores: one pointer add per point.  VAR_4A applies
        // W_L^{k*col} here: a table value every fourth point of a thread, the
        // step W_L^{TPT*col} in between (in the last pass it spilled).
        __syncthreads();
        if (b0 + c2 < a.count) {
            cx<T>* o = out + b0 + c2 + (long long)t2 * a.out_nd;
            const long long step = (long long)TPT * a.out_nd;
            const cx<T>* r2 = sm + c2 * N;
            const long long col = b0 + c2 + a.tw4_c0;
            cx<T> w = mk<T>(1, 0), wst = mk<T>(1, 0);
            if constexpr (VAR == VAR_4A) wst = tw4<T>(a, (long long)TPT * col);
            for (int k = t2, i = 0; k < N; k += TPT, ++i, o += step) {
                cx<T> y = r2[k];
                if constexpr (VAR == VAR_4A) {
                    if ((i & 3) == 0) w = tw4<T>(a, (long long)k * col);
                    else w = cmul(w, wst);
                    y = twmul<DIR>(y, w);
                }
                st_out(o, scl(y, scale), a.nt);
            }
        }
    }
}

}  // namespace afft
