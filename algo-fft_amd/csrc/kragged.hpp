// Persistent work-queue kernels for a ragged batch: thousands of independent
// 1-D complex transforms of different lengths (powers of two, 2/3/5/7-smooth
// and Bluestein sizes) in one queue per task type.
//
// The batch is cut into tasks.  Jobs whose FFT length m fits in one block
// (m <= rag_whole_cap: 16384 points complex64, 8192 complex128) run whole:
// power-of-two lengths from 2048 and Bluestein paddings from 2048 on the
// compile-time cores (one launch per length), every other length in one
// launch of the runtime passes.  A larger job is split four-step style,
// m = m1 * m2, into tasks that each hold a few whole sub-transforms in LDS:
//
//   column task   G adjacent columns n2 of x[m2*n1 + n2]: FFT over n1, times
//                 W_m^{n2*k1}, into S[k1*m2 + n2]
//   row task      G rows k1 of S: FFT over n2, stored as X[k1 + m1*k2]
//
// S is the job's own slice of the scratch array.  A Bluestein job's rows run
// the whole middle of the chain (forward FFT, filter product, inverse FFT,
// inverse twiddle W_m^{-k1*p}) in place in S, and a second set of column
// tasks finishes the inverse FFT over k1 and applies the post-chirp.
//
// Each task type runs as its own persistent kernel, with plain and Bluestein
// jobs and the transform direction as template parameters, so that every
// kernel holds only the one or two pass loops it needs (register pressure
// sets the occupancy).  A launch's blocks take the next task from the
// launch's counter until the counter passes the task count, which every
// block reaches.  Launch order: columns, rows, second columns, whole jobs.
// Tasks within a launch are sorted by decreasing cost on the host (a
// longest-processing-time queue).
//
// Passes run in place in one LDS buffer: each thread reads all of its
// butterflies of the pass into registers (at most ceil(16/R) of radix R,
// since a block holds at most 4096 points), the block passes a barrier, and
// the outputs are written back.  Global reads and writes go through staging
// loops whose lanes run along the contiguous dimension; sub-transforms sit in
// LDS rows of pitch L + 1 so column-wise staging is free of bank conflicts.
#pragma once
#include "kgen.hpp"

namespace afft {

constexpr int RAG_THREADS = 256;
constexpr int RAG_PTS = 4096;       // points per block (whole job, or G*L of a task)
constexpr int RAG_BUF = 4352;       // LDS buffer in points (G*(L+1) <= 4096 + G)
constexpr int RAG_WHOLE = 4096;     // largest sub-transform length of a split task
constexpr int RAG_SPASS = 12;       // passes of a schedule of length <= 16384

// Largest FFT length run as one task: the whole transform in 128 KiB of LDS.
constexpr int rag_whole_cap(int prec) { return prec ? 8192 : 16384; }

// Runtime-pass whole jobs: blocks of PTS points (the largest whole job) and
// PTS/16 threads; a task holds G = PTS/L jobs of one length L (a power of
// two), so every task moves about the same data and one launch serves every
// length.
template <typename T>
struct RagW {
    static constexpr int PTS = sizeof(T) == 4 ? 16384 : 8192;
    static constexpr int THREADS = PTS / 16;
};

struct RSched {
    int npass;
    int rad[RAG_SPASS];
    int ns[RAG_SPASS];
    unsigned nsm[RAG_SPASS];     // mod_magic(ns[p])
    int twoff[RAG_SPASS];
    const void* tw;              // per-pass twiddles W_{Ns*R}^{k*r}
};

enum { RT_WHOLE = 0, RT_COLS = 1, RT_ROWS = 2, RT_COLS2 = 3 };

// Launches, in order: columns of plain split jobs (generic passes, then the
// power-of-two cores), columns of Bluestein jobs, rows likewise, second
// columns of Bluestein jobs, whole plain jobs, whole Bluestein jobs.
enum {
    RL_COLS = 0,
    RL_COLS_P2,
    RL_COLS_BLUE,
    RL_ROWS,
    RL_ROWS_P2,
    RL_ROWS_BLUE,
    RL_COLS2,
    RL_WHOLE_P2,
    RL_WHOLE,
    RL_WHOLE_BLUE,
    RL_COUNT
};
constexpr int RAG_COUNTERS = 64;     // queue counters (one per launch) cleared by k_rag_reset
constexpr int RAG_P2_COL_MIN = 64, RAG_P2_COL_MAX = 1024;    // power-of-two column lengths m1
constexpr int RAG_P2_ROW_MIN = 128, RAG_P2_ROW_MAX = 2048;   // power-of-two row lengths m2

struct RJob {
    long long in_off, out_off;   // element offsets of the job's input and output
    long long scr_off;           // element offset of the job's S (split jobs)
    int n;                       // logical length
    int m;                       // FFT length (n, or the Bluestein padding)
    int m1, m2;                  // split jobs: m = m1 * m2
    int blue;                    // 1: Bluestein chain
    int g_col, g_row;            // sub-transforms per column / row task (powers of two)
    int shift;                   // W_m^i = hi[i >> shift] * lo[i & (2^shift - 1)]
    const void* twlo;
    const void* twhi;
    RSched s[2];                 // whole: s[0] of length m; split: s[0] of m1, s[1] of m2
    const void* chirp;           // Bluestein: n entries
    const void* filt[2];         // Bluestein: m entries each, [forward, inverse], 1/m folded in;
                                 // split jobs store them row-major as [k1][k2]
};

struct RTask {
    int job;
    int type;                    // RT_*
    int first;                   // first column / row
    int count;                   // columns / rows in the task (<= g_col / g_row)
};

struct RagArgs {
    const void* in;
    void* out;
    const RJob* jobs;
    const RTask* tasks;          // this launch's tasks
    int ntasks;
    int dir;
    int* counter;                // this launch's queue counter
    void* scratch;               // the split jobs' S arrays
    const int* joblist;          // grouped whole jobs: job ids, tasks index it
};

template <typename T>
__device__ __forceinline__ cx<T> rag_tw(const RJob& jb, long long i) {
    const gptr<cx<T>> lo = as_global<cx<T>>(jb.twlo);
    const gptr<cx<T>> hi = as_global<cx<T>>(jb.twhi);
    return cmul(lo[i & ((1LL << jb.shift) - 1)], hi[i >> jb.shift]);
}

// `cnt` transforms of length L in LDS rows of pitch P (transform c at
// buf + c*P), group c of TPT threads on transform c, all passes in place;
// the last pass hands its outputs to sink(c, k, v), which may write the
// transform's own row.  Every thread of the block calls this; the buffer is
// complete when it returns.  TPT is a power of two.
template <typename T, int DIR, class Sink>
__device__ __forceinline__ void rag_passes(const RSched& sc, int L, int P, int cnt, int TPT, cx<T>* buf, Sink sink) {
    const int tid = threadIdx.x;
    const int grp = tid >> (__ffs(TPT) - 1), t = tid & (TPT - 1);
    const bool active = grp < cnt;
    cx<T>* row = buf + grp * P;
    const gptr<cx<T>> tw = as_global<cx<T>>(sc.tw);
    for (int p = 0; p < sc.npass; ++p) {
        auto load = [&](int n) -> cx<T> { return row[n]; };
        auto store = [&](int k, cx<T> v) { row[k] = v; };
        auto fin = [&](int k, cx<T> v) { sink(grp, k, v); };
        const int R = sc.rad[p];
        if (p == sc.npass - 1)
            rt_pass_ip_any<DIR, T>(R, active, t, TPT, L, sc.ns[p], sc.nsm[p], tw + sc.twoff[p], load, fin);
        else
            rt_pass_ip_any<DIR, T>(R, active, t, TPT, L, sc.ns[p], sc.nsm[p], tw + sc.twoff[p], load, store);
        __syncthreads();
    }
}

// Staging between global memory and LDS.  A block moves at most RAG_PTS
// points, RAG_ITERS steps per thread.  rag_staged issues the global loads
// of a batch of steps before it uses the first one (about 128 bytes per
// thread in flight: 16 complex64 or 8 complex128 points, half that when a
// step loads two values), and no step needs an integer division.
constexpr int RAG_ITERS = RAG_PTS / RAG_THREADS;

// ld(s) returns step s's loaded value (its addresses must be valid for every
// s < nsteps: callers clamp them); put(s, v) consumes it, in order of s.
template <int BYTES, class Ld, class Put>
__device__ __forceinline__ void rag_staged(int nsteps, Ld ld, Put put) {
    constexpr int SB = 128 / BYTES < 1 ? 1 : (128 / BYTES > RAG_ITERS ? RAG_ITERS : 128 / BYTES);
#pragma unroll
    for (int h = 0; h < RAG_ITERS; h += SB) {
        if (h >= nsteps) break;
        decltype(ld(0)) v[SB];
#pragma unroll
        for (int s = 0; s < SB; ++s)
            if (h + s < nsteps) v[s] = ld(h + s);
#pragma unroll
        for (int s = 0; s < SB; ++s)
            if (h + s < nsteps) put(h + s, v[s]);
    }
}

template <typename T>
struct cx2 {
    cx<T> a, b;
};

// Tiles whose contiguous global dimension is the sub-transform index c
// (cnt <= G sub-transforms, G a power of two): lane c = tid mod G, line
// l = tid / G + s * (TH / G) in step s; nsteps(L) steps cover L lines.
template <int TH = RAG_THREADS>
struct ColLanes {
    int c, l0, step;
    __device__ explicit ColLanes(int G) {
        const int lg = __ffs(G) - 1;
        c = threadIdx.x & (G - 1);
        l0 = threadIdx.x >> lg;
        step = TH >> lg;
    }
    __device__ int line(int s) const { return l0 + s * step; }
    __device__ int nsteps(int L) const { return (L + step - 1) / step; }
};

// Flat index i = tid + s * RAG_THREADS of rows of length L, as (row c,
// element n), advanced by one step without a division.
template <int TH = RAG_THREADS>
struct RowWalk {
    int c, n, dq, dr, L;
    __device__ explicit RowWalk(int len) : L(len) {
        c = (int)threadIdx.x / L;
        n = (int)threadIdx.x - c * L;
        dq = TH / L;
        dr = TH - dq * L;
    }
    __device__ void next() {
        n += dr;
        c += dq;
        if (n >= L) {
            n -= L;
            ++c;
        }
    }
};

template <int TH = RAG_THREADS>
__device__ __forceinline__ int flat_steps(int total) { return (total + TH - 1) / TH; }

// Contiguous run src[0 .. total) into LDS rows of pitch L + 1 (row c of the
// run starts at src + c*L).
template <typename T, int TH = RAG_THREADS>
__device__ __forceinline__ void rag_stage_rows(const cx<T>* __restrict__ src, int total, int L, cx<T>* lds) {
    const int tid = threadIdx.x;
    RowWalk<TH> w(L);
    rag_staged<sizeof(cx<T>)>(
        flat_steps<TH>(total), [&](int s) { return src[min(tid + s * TH, total - 1)]; },
        [&](int s, cx<T> v) {
            const int i = tid + s * TH;
            if (i < total) lds[i + w.c] = v;
            w.next();
        });
    __syncthreads();
}

// Whole jobs of one length L on the runtime passes: a task holds up to
// G = jb.g_col of them (ids at a.joblist[first .. first+count)), one LDS row
// of pitch L + 1 and TH/G threads each (TH threads per block, at most 16
// points per thread).  The inputs are staged with contiguous lanes along
// each row; the last pass writes each job's output.
template <typename T, int DIR, int TH = RAG_THREADS>
__device__ void rag_whole(const RagArgs a, const RJob& j0, const RTask& tk, cx<T>* lds, long long* base) {
    const int L = j0.m, P = L + 1, TPT = TH / j0.g_col, cnt = tk.count, tid = threadIdx.x;
    const int total = cnt * L;
    if (tid < cnt) base[tid] = a.jobs[a.joblist[tk.first + tid]].in_off;
    __syncthreads();
    const cx<T>* __restrict__ x = (const cx<T>*)a.in;
    RowWalk<TH> rl(L), rw(L);
    rag_staged<sizeof(cx<T>)>(
        flat_steps<TH>(total),
        [&](int) {
            const cx<T> v = x[base[min(rl.c, cnt - 1)] + rl.n];
            rl.next();
            return v;
        },
        [&](int s, cx<T> v) {
            const int i = tid + s * TH;
            if (i < total) lds[i + rw.c] = v;
            rw.next();
        });
    __syncthreads();
    const int grp = tid / TPT;
    cx<T>* __restrict__ y = (cx<T>*)a.out + (grp < cnt ? a.jobs[a.joblist[tk.first + grp]].out_off : 0);
    const T sc = DIR > 0 ? T(1) / T(L) : T(1);
    auto sink = [&](int, int k, cx<T> v) { y[k] = scl(v, sc); };
    rag_passes<T, DIR>(j0.s[0], L, P, cnt, TPT, lds, sink);
}

// A whole Bluestein job on the runtime passes (one job per block of
// RAG_THREADS threads, m <= RAG_PTS): pre-chirp and zero padding on the
// staging load, forward FFT, filter product, inverse FFT, post-chirp on the
// store.
template <typename T>
__device__ void rag_whole_blue(const RagArgs a, const RJob& jb, cx<T>* lds) {
    const cx<T>* __restrict__ x = (const cx<T>*)a.in + jb.in_off;
    cx<T>* __restrict__ y = (cx<T>*)a.out + jb.out_off;
    const int m = jb.m, n = jb.n, tid = threadIdx.x;
    const bool inv = a.dir > 0;
    const gptr<cx<T>> chirp = as_global<cx<T>>(jb.chirp);
    const gptr<cx<T>> filt = as_global<cx<T>>(jb.filt[inv ? 1 : 0]);
    rag_staged<sizeof(cx2<T>)>(
        flat_steps(m),
        [&](int s) {
            const int i = min(tid + s * RAG_THREADS, n - 1);
            return cx2<T>{x[i], chirp[i]};
        },
        [&](int s, cx2<T> v) {
            const int i = tid + s * RAG_THREADS;
            if (i < m) lds[i] = i < n ? (inv ? cmulc(v.a, v.b) : cmul(v.a, v.b)) : mk<T>(0, 0);
        });
    __syncthreads();
    const T sc = inv ? T(1) / T(n) : T(1);
    auto sink = [&](int, int k, cx<T> v) { lds[k] = v; };
    rag_passes<T, -1>(jb.s[0], m, m, 1, RAG_THREADS, lds, sink);
    rag_staged<sizeof(cx<T>)>(
        flat_steps(m), [&](int s) { return filt[min(tid + s * RAG_THREADS, m - 1)]; },
        [&](int s, cx<T> w) {
            const int i = tid + s * RAG_THREADS;
            if (i < m) lds[i] = cmul(lds[i], w);
        });
    __syncthreads();
    rag_passes<T, 1>(jb.s[0], m, m, 1, RAG_THREADS, lds, sink);
    rag_staged<sizeof(cx<T>)>(
        flat_steps(n), [&](int s) { return chirp[min(tid + s * RAG_THREADS, n - 1)]; },
        [&](int s, cx<T> w) {
            const int i = tid + s * RAG_THREADS;
            if (i < n) y[i] = scl(inv ? cmulc(lds[i], w) : cmul(lds[i], w), sc);
        });
}

// Tasks of the generic splits: at most PTS points (G <= THREADS
// sub-transforms of L points, G*L <= PTS) on THREADS threads, 16 points per
// thread at most, in LDS rows of pitch L + 1.  Complex64 tasks of 8192
// points on 512 threads held to 128 VGPRs (4 waves per SIMD, 7-12 VGPRs
// spilled) read twice the adjacent columns per row of a 4096-point task.
// Measured on the smooth split jobs of a 4096-job mixed set: 0.146 ms with
// 4096-point tasks (161-164 VGPRs, 3 waves per SIMD), 0.149 ms with those
// held to 128 VGPRs, 0.162 ms with 8192-point tasks at 165 VGPRs (one block
// per CU), 0.131 ms with them at 128, 0.139 ms with 16384-point tasks.
template <typename T>
struct RagG {
    static constexpr int PTS = sizeof(T) == 4 ? 8192 : RAG_PTS;
    static constexpr int THREADS = PTS / 16;
    static constexpr int BUF = PTS + THREADS;
    static constexpr int MIN_WAVES = sizeof(T) == 4 ? 4 : 2;   // per SIMD
};

// Column task of a split job with generic passes: columns c0 .. c0+cnt-1 of
// the m1 x m2 view x[m2*n1 + n2], FFT over n1 with sign SD, times
// W_m^{SD * n2 * k1}, into S[k1*m2 + n2].
template <typename T, int SD, int TH>
__device__ void rag_cols(const RagArgs a, const RJob& jb, const RTask& tk, cx<T>* lds) {
    const int L = jb.m1, P = L + 1, G = jb.g_col, TPT = TH / G;
    const int cnt = tk.count, c0 = tk.first, m2 = jb.m2;
    const cx<T>* __restrict__ x = (const cx<T>*)a.in + jb.in_off;
    cx<T>* __restrict__ S = (cx<T>*)a.scratch + jb.scr_off;
    const ColLanes<TH> ln(G);
    const bool lane = ln.c < cnt;
    const int cc = c0 + min(ln.c, cnt - 1);
    const int ns = ln.nsteps(L);
    rag_staged<sizeof(cx<T>)>(
        ns, [&](int s) { return x[(long long)m2 * min(ln.line(s), L - 1) + cc]; },
        [&](int s, cx<T> v) {
            const int n1 = ln.line(s);
            if (n1 < L && lane) lds[ln.c * P + n1] = v;
        });
    __syncthreads();
    auto sink = [&](int c, int k1, cx<T> v) { lds[c * P + k1] = v; };
    rag_passes<T, SD>(jb.s[0], L, P, cnt, TPT, lds, sink);
    // W_m^{cc*k1} for k1 = line(s): a table value every fourth step and the
    // step W_m^{cc*lines per step} in between (at most three products).
    // Against a two-level table lookup per point: the smooth split jobs of a
    // 4096-job mixed set 0.130 -> 0.123 ms.
    const cx<T> st = rag_tw<T>(jb, (long long)cc * ln.step);
    cx<T> w = mk<T>(1, 0);
#pragma unroll
    for (int s = 0; s < RAG_ITERS; ++s) {
        if (s >= ns) break;
        const int k1 = ln.line(s);
        w = s % 4 == 0 ? rag_tw<T>(jb, (long long)cc * min(k1, L - 1)) : cmul(w, st);
        if (k1 < L && lane) S[(long long)k1 * m2 + cc] = twmul<SD>(lds[ln.c * P + k1], w);
    }
}

// Row task of a split job: rows r0 .. r0+cnt-1 of S (length m2), FFT over n2
// with sign SD, output X[k1 + m1*k2] (scaled 1/m for the inverse).
template <typename T, int SD, int TH>
__device__ void rag_rows(const RagArgs a, const RJob& jb, const RTask& tk, cx<T>* lds) {
    const int L = jb.m2, P = L + 1, G = jb.g_row, TPT = TH / G;
    const int cnt = tk.count, r0 = tk.first, m1 = jb.m1;
    rag_stage_rows<T, TH>((const cx<T>*)a.scratch + jb.scr_off + (long long)r0 * L, cnt * L, L, lds);
    auto sink = [&](int c, int k2, cx<T> v) { lds[c * P + k2] = v; };
    rag_passes<T, SD>(jb.s[1], L, P, cnt, TPT, lds, sink);
    const T sc = SD > 0 ? T(1) / T(jb.m) : T(1);
    cx<T>* __restrict__ y = (cx<T>*)a.out + jb.out_off + r0;
    const ColLanes<TH> ln(G);
    if (ln.c < cnt) {
        const int ns = ln.nsteps(L);
#pragma unroll
        for (int s = 0; s < RAG_ITERS; ++s) {
            if (s >= ns) break;
            const int k2 = ln.line(s);
            if (k2 < L) y[ln.c + (long long)m1 * k2] = scl(lds[ln.c * P + k2], sc);
        }
    }
}

// ---- power-of-two sub-transforms on the compile-time cores -------------------
// Split jobs whose sub-transform lengths are powers of two (plain
// power-of-two jobs above the whole-job cap, and every split Bluestein job,
// whose m is its power-of-two padding; m <= 2^21 gives m1 in [64, 1024], m2
// in [128, 2048])
// run their column and row tasks on the radix-16 register cores of the
// single-size kernels: 16 points per thread, L/16 threads per sub-transform,
// G = RagS<T>::PTS/L sub-transforms per task (the host cuts every such job
// into tasks of exactly G), global loads straight into registers.  Each length L
// is its own kernel and launch: one kernel holding several lengths would need
// the registers of all of them.  The twiddles come from the job's runtime
// schedule, whose table has the layout of the matching Sched (the remainder
// radix first, no twiddles in pass 0).

template <int L>
struct P2Sched;
template <> struct P2Sched<16> { using S = Sched<16>; };
template <> struct P2Sched<32> { using S = Sched<2, 16>; };
template <> struct P2Sched<64> { using S = Sched<4, 16>; };
template <> struct P2Sched<128> { using S = Sched<8, 16>; };
template <> struct P2Sched<256> { using S = Sched<16, 16>; };
template <> struct P2Sched<512> { using S = Sched<2, 16, 16>; };
template <> struct P2Sched<1024> { using S = Sched<4, 16, 16>; };
template <> struct P2Sched<2048> { using S = Sched<8, 16, 16>; };
template <> struct P2Sched<4096> { using S = Sched<16, 16, 16>; };
template <> struct P2Sched<8192> { using S = Sched<2, 16, 16, 16>; };
template <> struct P2Sched<16384> { using S = Sched<4, 16, 16, 16>; };

// Tasks of the power-of-two splits: PTS points (G = PTS/L sub-transforms of
// L points) on PTS/16 threads.  8192 points for complex64 double the
// adjacent columns a column task reads per row (256 bytes at L = 256)
// against 4096-point tasks: job offsets in the packed layout are not
// aligned, so each row segment also touches a neighbouring cache line.
// Measured on a 4096-job mixed set: HBM traffic 1.59 -> 1.50 GB per
// execution at the same graph time.
template <typename T>
struct RagS {
    static constexpr int PTS = sizeof(T) == 4 ? 8192 : 4096;
    static constexpr int THREADS = PTS / 16;
};

// LDS of the row tasks: G padded slabs (the column tasks need PTS points).
template <typename T>
struct RagP2 {
    static constexpr int BUF = RagS<T>::PTS + (RagS<T>::PTS >> PadFor<T, 64>::value) + RagS<T>::PTS / 64;
};

// Column task: columns c0 .. c0+G-1 of x[m2*n1 + n2] (chirp * x zero-padded
// past n for Bluestein jobs), FFT over n1 with sign SD, times
// W_m^{SD * n2 * k1}, into S[k1*m2 + n2].
template <typename T, int L, int SD, bool BLUE>
__device__ __forceinline__ void p2_cols(const RagArgs a, const RJob& jb, const RTask& tk, cx<T>* lds) {
    using S = typename P2Sched<L>::S;
    constexpr int E = 16, TT = L / E, G = RagS<T>::PTS / L;
    const int tid = threadIdx.x, c = tid % G, t = tid / G;
    const long long col = tk.first + c, m2 = jb.m2;
    const cx<T>* __restrict__ x = (const cx<T>*)a.in + jb.in_off;
    cx<T> v[E];
    if constexpr (BLUE) {
        const gptr<cx<T>> chirp = as_global<cx<T>>(jb.chirp);
        const bool inv = a.dir > 0;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const long long idx = m2 * (t + TT * e) + col;
            v[e] = mk<T>(0, 0);
            if (idx < jb.n) v[e] = inv ? cmulc(x[idx], chirp[idx]) : cmul(x[idx], chirp[idx]);
        }
    } else {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = x[m2 * (t + TT * e) + col];
    }
    core<T, S, E, SD>(v, lds, t, as_global<cx<T>>(jb.s[0].tw), LdsCol<G>(c));
    cx<T>* __restrict__ sp = (cx<T>*)a.scratch + jb.scr_off + col;
    // W_m^{k1*col} for k1 = t + TT*e: a table value every fourth e, the step
    // W_m^{TT*col} in between (lanes differ in col: scattered table lines)
    const cx<T> st = rag_tw<T>(jb, col * TT);
#pragma unroll
    for (int e0 = 0; e0 < E; e0 += 4) {
        cx<T> w = rag_tw<T>(jb, col * (t + TT * e0));
#pragma unroll
        for (int e = e0; e < e0 + 4; ++e) {
            if (e > e0) w = cmul(w, st);
            sp[m2 * (t + TT * e)] = twmul<SD>(v[e], w);
        }
    }
}

// The G rows r0 .. r0+G-1 of S, one per group of L/16 threads.
template <typename T, int L>
__device__ __forceinline__ void p2_row_load(const cx<T>* __restrict__ row, int t, cx<T> (&v)[16]) {
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = row[t + (L / 16) * e];
}

// Row task: rows r0 .. r0+G-1 of S, FFT over n2 with sign SD, output
// X[k1 + m1*k2] (scaled 1/m for the inverse), stored through LDS so that the
// lanes of a store run along k1.
template <typename T, int L, int SD>
__device__ __forceinline__ void p2_rows(const RagArgs a, const RJob& jb, const RTask& tk, cx<T>* lds) {
    using S = typename P2Sched<L>::S;
    constexpr int E = 16, TT = L / E, G = RagS<T>::PTS / L, PADSH = PadFor<T, L>::value;
    const int tid = threadIdx.x, c = tid / TT, t = tid % TT;
    const int r0 = tk.first;
    cx<T> v[E];
    p2_row_load<T, L>((const cx<T>*)a.scratch + jb.scr_off + (long long)(r0 + c) * L, t, v);
    const LdsRow<L, PADSH> li(c);
    core<T, S, E, SD>(v, lds, t, as_global<cx<T>>(jb.s[1].tw), li);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) lds[li(t + TT * e)] = v[e];
    __syncthreads();
    const int c2 = tid % G, t2 = tid / G;
    const LdsRow<L, PADSH> li2(c2);
    const T sc = SD > 0 ? T(1) / T(jb.m) : T(1);
    cx<T>* __restrict__ y = (cx<T>*)a.out + jb.out_off + r0 + c2;
    const long long m1 = jb.m1;
#pragma unroll
    for (int e = 0; e < E; ++e) y[m1 * (t2 + TT * e)] = scl(lds[li2(t2 + TT * e)], sc);
}

// Row task of a Bluestein job: forward FFT, filter product (the filter is
// stored row-major as [k1][k2]), inverse FFT, times W_m^{-k1*p}, back into
// the same rows of S.
template <typename T, int L>
__device__ __forceinline__ void p2_rows_blue(const RagArgs a, const RJob& jb, const RTask& tk, cx<T>* lds) {
    using S = typename P2Sched<L>::S;
    constexpr int E = 16, TT = L / E, PADSH = PadFor<T, L>::value;
    const int tid = threadIdx.x, c = tid / TT, t = tid % TT;
    const int k1 = tk.first + c;
    cx<T>* __restrict__ row = (cx<T>*)a.scratch + jb.scr_off + (long long)k1 * L;
    const gptr<cx<T>> frow = as_global<cx<T>>(jb.filt[a.dir > 0 ? 1 : 0]) + (long long)k1 * L;
    const gptr<cx<T>> tw = as_global<cx<T>>(jb.s[1].tw);
    cx<T> v[E];
    p2_row_load<T, L>(row, t, v);
    const LdsRow<L, PADSH> li(c);
    core<T, S, E, -1>(v, lds, t, tw, li);
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = cmul(v[e], frow[t + TT * e]);
    core<T, S, E, 1, true>(v, lds, t, tw, li);
    // W_m^{k1*p} for p = t + TT*e: a table value every fourth e, the step
    // W_m^{k1*TT} in between (the split Bluestein jobs of a 4096-job mixed
    // set: 0.155 -> 0.145 ms)
    const cx<T> st = rag_tw<T>(jb, (long long)k1 * TT);
    cx<T> w = mk<T>(1, 0);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int p = t + TT * e;
        w = e % 4 == 0 ? rag_tw<T>(jb, (long long)k1 * p) : cmul(w, st);
        row[p] = twmul<1>(v[e], w);
    }
}

// Whole power-of-two jobs of one length L: a task holds up to
// G = max(1, RAG_PTS/L) of them (their ids at a.joblist[first .. first+count)),
// L/16 threads each, loaded straight into registers and stored in natural
// order.  Lengths above RAG_PTS run one job per block of L/16 threads.  The
// table pointer is the first job's: every job of the task has the same
// schedule.
template <typename T, int L>
struct P2Whole {
    static constexpr int PADSH = PadFor<T, L>::value;
    static constexpr int G = L >= RAG_PTS ? 1 : RAG_PTS / L;
    static constexpr int THREADS = G * L / 16;
    static constexpr int LDS = G * LdsRow<L, PADSH>::SLAB;
};

template <typename T, int L, int DIR>
__device__ __forceinline__ void p2_whole(const RagArgs a, const RTask& tk, cx<T>* lds) {
    using S = typename P2Sched<L>::S;
    constexpr int E = 16, TT = L / E;
    const int tid = threadIdx.x, c = tid / TT, t = tid % TT;
    const bool act = c < tk.count;
    const int jid = a.joblist[tk.first + (act ? c : 0)];
    const cx<T>* __restrict__ x = (const cx<T>*)a.in + a.jobs[jid].in_off;
    const gptr<cx<T>> tw = as_global<cx<T>>(a.jobs[a.joblist[tk.first]].s[0].tw);
    cx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = x[t + TT * e];
    core<T, S, E, DIR>(v, lds, t, tw, LdsRow<L, P2Whole<T, L>::PADSH>(c));
    if (act) {
        const T sc = DIR > 0 ? T(1) / T(L) : T(1);
        cx<T>* __restrict__ y = (cx<T>*)a.out + a.jobs[jid].out_off;
#pragma unroll
        for (int e = 0; e < E; ++e) y[t + TT * e] = scl(v[e], sc);
    }
}

// Whole Bluestein jobs of one padded length L (grouped like p2_whole; the
// jobs of a task differ in n, so each group reads its own job's chirp and
// filter): pre-chirp and zero padding on the load, forward FFT, filter
// product, inverse FFT, post-chirp (and 1/n for the inverse transform) on
// the store, all on the compile-time core.
// Points per thread of the Bluestein whole-job cores (16: L/16 threads per
// job; 32 points a thread at 16384 points measured worse at compile time,
// 256 registers and more spills than the 128-register 1024-thread form).
template <typename T, int L>
struct P2WholeBlue {
    static constexpr int E = 16;
    static constexpr int THREADS = P2Whole<T, L>::G * L / E;
};

template <typename T, int L>
__device__ __forceinline__ void p2_whole_blue(const RagArgs a, const RTask& tk, cx<T>* lds) {
    using S = typename P2Sched<L>::S;
    constexpr int E = P2WholeBlue<T, L>::E, TT = L / E;
    const int tid = threadIdx.x, c = tid / TT, t = tid % TT;
    const bool act = c < tk.count;
    const RJob& jb = a.jobs[a.joblist[tk.first + (act ? c : 0)]];
    const bool inv = a.dir > 0;
    const int n = jb.n;
    const cx<T>* __restrict__ x = (const cx<T>*)a.in + jb.in_off;
    const gptr<cx<T>> chirp = as_global<cx<T>>(jb.chirp);
    const gptr<cx<T>> filt = as_global<cx<T>>(jb.filt[inv ? 1 : 0]);
    const gptr<cx<T>> tw = as_global<cx<T>>(a.jobs[a.joblist[tk.first]].s[0].tw);
    const LdsRow<L, P2Whole<T, L>::PADSH> li(c);
    cx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = t + TT * e;
        v[e] = mk<T>(0, 0);
        if (i < n) v[e] = inv ? cmulc(x[i], chirp[i]) : cmul(x[i], chirp[i]);
    }
    core<T, S, E, -1>(v, lds, t, tw, li);
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = cmul(v[e], filt[t + TT * e]);
    core<T, S, E, 1, true>(v, lds, t, tw, li);
    if (act) {
        const T sc = inv ? T(1) / T(n) : T(1);
        cx<T>* __restrict__ y = (cx<T>*)a.out + jb.out_off;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int i = t + TT * e;
            if (i < n) y[i] = scl(inv ? cmulc(v[e], chirp[i]) : cmul(v[e], chirp[i]), sc);
        }
    }
}

// Second column task of a Bluestein job: columns p0 .. p0+G-1 of S, inverse
// FFT over k1, output z[p + m2*q] times the post-chirp (and 1/n for the
// inverse transform), cropped to n.
template <typename T, int L>
__device__ __forceinline__ void p2_cols2(const RagArgs a, const RJob& jb, const RTask& tk, cx<T>* lds) {
    using S = typename P2Sched<L>::S;
    constexpr int E = 16, TT = L / E, G = RagS<T>::PTS / L;
    const int tid = threadIdx.x, c = tid % G, t = tid / G;
    const long long p = tk.first + c, m2 = jb.m2;
    const cx<T>* __restrict__ sp = (const cx<T>*)a.scratch + jb.scr_off + p;
    cx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = sp[m2 * (t + TT * e)];
    core<T, S, E, 1>(v, lds, t, as_global<cx<T>>(jb.s[0].tw), LdsCol<G>(c));
    const bool inv = a.dir > 0;
    const T sc = inv ? T(1) / T(jb.n) : T(1);
    const gptr<cx<T>> chirp = as_global<cx<T>>(jb.chirp);
    cx<T>* __restrict__ y = (cx<T>*)a.out + jb.out_off;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const long long n = p + m2 * (t + TT * e);
        if (n < jb.n) y[n] = scl(inv ? cmulc(v[e], chirp[n]) : cmul(v[e], chirp[n]), sc);
    }
}

// The persistent loop shared by the task kernels.
template <class Run>
__device__ __forceinline__ void rag_queue(const RagArgs a, Run run) {
    __shared__ int s_task;
    const int tid = threadIdx.x;
    while (true) {
        if (tid == 0) s_task = atomicAdd(a.counter, 1);
        __syncthreads();
        const int id = s_task;
        __syncthreads();
        if (id >= a.ntasks) break;
        const RTask tk = a.tasks[id];
        run(a.jobs[tk.job], tk);
    }
}

// The kernels of the launches (RL_* in launch order).
template <typename T, int SD>
__global__ void __launch_bounds__(RagG<T>::THREADS, RagG<T>::MIN_WAVES) k_rag_cols(RagArgs a) {
    __shared__ cx<T> lds[RagG<T>::BUF];
    rag_queue(a, [=](const RJob& jb, const RTask& tk) { rag_cols<T, SD, RagG<T>::THREADS>(a, jb, tk, lds); });
}

template <typename T, int L, int SD, bool BLUE>
__global__ void __launch_bounds__(RagS<T>::THREADS) k_rag_cols_p2(RagArgs a) {
    __shared__ cx<T> lds[RagS<T>::PTS];
    rag_queue(a, [=](const RJob& jb, const RTask& tk) { p2_cols<T, L, SD, BLUE>(a, jb, tk, lds); });
}

template <typename T, int SD>
__global__ void __launch_bounds__(RagG<T>::THREADS, RagG<T>::MIN_WAVES) k_rag_rows(RagArgs a) {
    __shared__ cx<T> lds[RagG<T>::BUF];
    rag_queue(a, [=](const RJob& jb, const RTask& tk) { rag_rows<T, SD, RagG<T>::THREADS>(a, jb, tk, lds); });
}

template <typename T, int L, int SD>
__global__ void __launch_bounds__(RagS<T>::THREADS) k_rag_rows_p2(RagArgs a) {
    __shared__ cx<T> lds[RagP2<T>::BUF];
    rag_queue(a, [=](const RJob& jb, const RTask& tk) { p2_rows<T, L, SD>(a, jb, tk, lds); });
}

template <typename T, int L>
__global__ void __launch_bounds__(RagS<T>::THREADS) k_rag_rows_blue(RagArgs a) {
    __shared__ cx<T> lds[RagP2<T>::BUF];
    rag_queue(a, [=](const RJob& jb, const RTask& tk) { p2_rows_blue<T, L>(a, jb, tk, lds); });
}

template <typename T, int L>
__global__ void __launch_bounds__(RagS<T>::THREADS) k_rag_cols2(RagArgs a) {
    __shared__ cx<T> lds[RagS<T>::PTS];
    rag_queue(a, [=](const RJob& jb, const RTask& tk) { p2_cols2<T, L>(a, jb, tk, lds); });
}

// The whole-job kernels.  One kernel per length L for the compile-time
// cores: a kernel holding several lengths needs the registers of all of
// them (and spills at 1024 threads per block), so lengths below 2048, whose
// jobs hold little of the data, run on the runtime passes instead, all
// lengths in one launch.
template <typename T, int L, int DIR>
__global__ void __launch_bounds__((P2Whole<T, L>::THREADS)) k_rag_whole_p2(RagArgs a) {
    __shared__ cx<T> lds[P2Whole<T, L>::LDS];
    rag_queue(a, [=](const RJob&, const RTask& tk) { p2_whole<T, L, DIR>(a, tk, lds); });
}

template <typename T, int L>
__global__ void __launch_bounds__((P2WholeBlue<T, L>::THREADS)) k_rag_whole_blue_p2(RagArgs a) {
    __shared__ cx<T> lds[P2Whole<T, L>::LDS];
    rag_queue(a, [=](const RJob&, const RTask& tk) { p2_whole_blue<T, L>(a, tk, lds); });
}

template <typename T>
__global__ void __launch_bounds__(RAG_THREADS) k_rag_whole_blue(RagArgs a) {
    __shared__ cx<T> lds[RAG_BUF];
    rag_queue(a, [=](const RJob& jb, const RTask&) { rag_whole_blue<T>(a, jb, lds); });
}

template <typename T, int DIR>
__global__ void __launch_bounds__(RagW<T>::THREADS) k_rag_whole(RagArgs a) {
    __shared__ cx<T> lds[RagW<T>::PTS + RagW<T>::THREADS];
    __shared__ long long base[RagW<T>::THREADS];
    rag_queue(a, [=](const RJob& jb, const RTask& tk) {
        rag_whole<T, DIR, RagW<T>::THREADS>(a, jb, tk, lds, base);
    });
}


}  // namespace afft
