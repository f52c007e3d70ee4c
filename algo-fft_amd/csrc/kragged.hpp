// Persistent work-queue kernel for a ragged batch: thousands of independent
// 1-D complex transforms of different lengths (powers of two, 2/3/5/7-smooth
// and Bluestein sizes) in one queue.
//
// The batch is cut into tasks.  A job whose FFT length m fits in one block
// (m <= RAG_WHOLE) is one task, run with the runtime Stockham passes of
// kgen.hpp between two LDS buffers.  A larger job is split four-step style,
// m = m1 * m2, into tasks that each hold a few whole sub-transforms in LDS:
//
//   column task  (phase 0)  G adjacent columns n2 of x[m2*n1 + n2]: FFT over
//                           n1, times W_m^{n2*k1}, into S[k1*m2 + n2]
//   row task     (phase 1)  G rows k1 of S: FFT over n2, stored as
//                           X[k1 + m1*k2]
//
// S is the job's own slice of the scratch array.  A Bluestein job's rows run
// the whole middle of the chain (forward FFT, filter product, inverse FFT,
// inverse twiddle W_m^{-k1*p}) in place in S, and a third pass of column tasks
// (phase 2) finishes the inverse FFT over k1 and applies the post-chirp.
//
// Each phase is one launch of the same persistent kernel: every block takes
// the next task from the phase's counter until the counter passes the task
// count, which every block reaches.  Tasks within a phase are sorted by
// decreasing cost on the host (a longest-processing-time queue).
//
// Global reads and writes of the column and row tasks go through an LDS
// staging loop whose lanes run along the contiguous dimension; sub-transforms
// sit in LDS rows of pitch L + 1 so column-wise staging is free of bank
// conflicts.
#pragma once
#include "kgen.hpp"

namespace afft {

constexpr int RAG_THREADS = 256;
constexpr int RAG_HALF = 4352;      // complex points per LDS buffer (two buffers)
constexpr int RAG_WHOLE = 4096;     // largest FFT length run as a single task
constexpr int RAG_SPASS = 12;       // passes of a schedule of length <= 4096

struct RSched {
    int npass;
    int rad[RAG_SPASS];
    int ns[RAG_SPASS];
    int twoff[RAG_SPASS];
    const void* tw;              // per-pass twiddles W_{Ns*R}^{k*r}
};

enum { RT_WHOLE = 0, RT_COLS = 1, RT_ROWS = 2, RT_COLS2 = 3 };

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
    const RTask* tasks;          // this phase's tasks
    int ntasks;
    int dir;
    int* counter;                // this phase's queue counter
    void* scratch;               // the split jobs' S arrays
};

template <typename T>
__device__ __forceinline__ cx<T>* rag_fin(int npass, cx<T>* A, cx<T>* B) {
    return ((npass - 1) & 1) ? A : B;
}

template <typename T>
__device__ __forceinline__ cx<T> rag_tw(const RJob& jb, long long i) {
    const cx<T>* lo = (const cx<T>*)jb.twlo;
    const cx<T>* hi = (const cx<T>*)jb.twhi;
    return cmul(lo[i & ((1LL << jb.shift) - 1)], hi[i >> jb.shift]);
}

// `cnt` transforms of length L, transform c in LDS row c (pitch P), group c
// of TPT threads on transform c.  Pass p reads buffer (p & 1 ? B : A) and
// writes the other; the last pass hands its outputs to sink(c, k, v), which
// may write rag_fin(npass, A, B).  Every thread of the block calls this (one
// barrier per pass).
template <typename T, int DIR, class Sink>
__device__ __forceinline__ void rag_passes(const RSched& sc, int L, int P, int cnt, int TPT, cx<T>* A, cx<T>* B,
                                           Sink sink) {
    const int tid = threadIdx.x;
    const int grp = tid / TPT, t = tid - grp * TPT;
    const cx<T>* __restrict__ tw = (const cx<T>*)sc.tw;
    for (int p = 0; p < sc.npass; ++p) {
        const bool last = p == sc.npass - 1;
        cx<T>* rd = ((p & 1) ? B : A) + grp * P;
        cx<T>* wr = ((p & 1) ? A : B) + grp * P;
        if (grp < cnt) {
            auto load = [&](int n) -> cx<T> { return rd[n]; };
            auto store = [&](int k, cx<T> v) { wr[k] = v; };
            auto fin = [&](int k, cx<T> v) { sink(grp, k, v); };
            const int R = sc.rad[p];
            if (R == 1) {
                for (int j = t; j < L; j += TPT) fin(j, load(j));
            } else if (last) {
                rt_pass_any<DIR>(R, t, TPT, L, sc.ns[p], tw + sc.twoff[p], load, fin);
            } else {
                rt_pass_any<DIR>(R, t, TPT, L, sc.ns[p], tw + sc.twoff[p], load, store);
            }
        }
        __syncthreads();
    }
}

// A whole job in one block: x staged into LDS with contiguous lanes, the
// passes in LDS, the last pass writing y (its output index runs contiguously
// across the threads as well).
template <typename T>
__device__ void rag_whole(const RagArgs a, const RJob& jb, cx<T>* lds) {
    const cx<T>* __restrict__ x = (const cx<T>*)a.in + jb.in_off;
    cx<T>* __restrict__ y = (cx<T>*)a.out + jb.out_off;
    const int m = jb.m, tid = threadIdx.x;
    cx<T>* A = lds;
    cx<T>* B = lds + RAG_HALF;
    const bool inv = a.dir > 0;
    if (!jb.blue) {
        for (int i = tid; i < m; i += RAG_THREADS) A[i] = x[i];
        __syncthreads();
        const T sc = inv ? T(1) / T(m) : T(1);
        auto sink = [&](int, int k, cx<T> v) { y[k] = scl(v, sc); };
        if (inv) rag_passes<T, 1>(jb.s[0], m, m, 1, RAG_THREADS, A, B, sink);
        else rag_passes<T, -1>(jb.s[0], m, m, 1, RAG_THREADS, A, B, sink);
        return;
    }
    // pre-chirp and zero padding on the staging load, filter product on the
    // forward FFT's last store (into C, the buffer its last pass does not
    // read), the inverse FFT from C, post-chirp on its last store
    const cx<T>* __restrict__ chirp = (const cx<T>*)jb.chirp;
    const cx<T>* __restrict__ filt = (const cx<T>*)jb.filt[inv ? 1 : 0];
    const int n = jb.n;
    for (int i = tid; i < m; i += RAG_THREADS) {
        cx<T> v = mk<T>(0, 0);
        if (i < n) v = inv ? cmulc(x[i], chirp[i]) : cmul(x[i], chirp[i]);
        A[i] = v;
    }
    __syncthreads();
    const T sc = inv ? T(1) / T(n) : T(1);
    cx<T>* C = rag_fin(jb.s[0].npass, A, B);
    cx<T>* D = C == A ? B : A;
    auto sink1 = [&](int, int k, cx<T> v) { C[k] = cmul(v, filt[k]); };
    rag_passes<T, -1>(jb.s[0], m, m, 1, RAG_THREADS, A, B, sink1);
    auto sink2 = [&](int, int k, cx<T> v) {
        if (k < n) y[k] = scl(inv ? cmulc(v, chirp[k]) : cmul(v, chirp[k]), sc);
    };
    rag_passes<T, 1>(jb.s[0], m, m, 1, RAG_THREADS, C, D, sink2);
}

// Column task of a split job (phase 0): columns c0 .. c0+cnt-1 of the m1 x m2
// view x[m2*n1 + n2], FFT over n1 with sign SD, times W_m^{SD * n2 * k1}, into
// S[k1*m2 + n2].  Bluestein jobs load chirp * x, zero past n.
template <typename T, int SD>
__device__ void rag_cols(const RagArgs a, const RJob& jb, const RTask& tk, cx<T>* lds) {
    const int L = jb.m1, P = L + 1, TPT = RAG_THREADS / jb.g_col;
    const int cnt = tk.count, c0 = tk.first, m2 = jb.m2, tid = threadIdx.x;
    const cx<T>* __restrict__ x = (const cx<T>*)a.in + jb.in_off;
    const cx<T>* __restrict__ chirp = (const cx<T>*)jb.chirp;
    cx<T>* __restrict__ S = (cx<T>*)a.scratch + jb.scr_off;
    cx<T>* A = lds;
    cx<T>* B = lds + RAG_HALF;
    const bool inv = a.dir > 0;
    for (int i = tid; i < cnt * L; i += RAG_THREADS) {
        const int n1 = i / cnt, c = i - n1 * cnt;
        const long long idx = (long long)m2 * n1 + c0 + c;
        cx<T> v;
        if (jb.blue) {
            v = mk<T>(0, 0);
            if (idx < jb.n) v = inv ? cmulc(x[idx], chirp[idx]) : cmul(x[idx], chirp[idx]);
        } else {
            v = x[idx];
        }
        A[c * P + n1] = v;
    }
    __syncthreads();
    cx<T>* F = rag_fin(jb.s[0].npass, A, B);
    auto sink = [&](int c, int k1, cx<T> v) {
        F[c * P + k1] = twmul<SD>(v, rag_tw<T>(jb, (long long)(c0 + c) * k1));
    };
    rag_passes<T, SD>(jb.s[0], L, P, cnt, TPT, A, B, sink);
    for (int i = tid; i < cnt * L; i += RAG_THREADS) {
        const int k1 = i / cnt, c = i - k1 * cnt;
        S[(long long)k1 * m2 + c0 + c] = F[c * P + k1];
    }
}

// Row task of a split job (phase 1): rows r0 .. r0+cnt-1 of S (length m2).
// Plain jobs: FFT over n2 with sign SD, output X[k1 + m1*k2] (scaled 1/m for
// the inverse).  Bluestein jobs: forward FFT, filter product, inverse FFT,
// times W_m^{-k1*p}, back into the same row of S.
template <typename T, int SD>
__device__ void rag_rows(const RagArgs a, const RJob& jb, const RTask& tk, cx<T>* lds) {
    const int L = jb.m2, P = L + 1, TPT = RAG_THREADS / jb.g_row;
    const int cnt = tk.count, r0 = tk.first, m1 = jb.m1, tid = threadIdx.x;
    cx<T>* __restrict__ S = (cx<T>*)a.scratch + jb.scr_off;
    cx<T>* A = lds;
    cx<T>* B = lds + RAG_HALF;
    for (int i = tid; i < cnt * L; i += RAG_THREADS) {
        const int c = i / L, n2 = i - c * L;
        A[c * P + n2] = S[(long long)(r0 + c) * L + n2];
    }
    __syncthreads();
    if (!jb.blue) {
        cx<T>* F = rag_fin(jb.s[1].npass, A, B);
            auto sink = [&](int c, int k2, cx<T> v) { F[c * P + k2] = v; };
        rag_passes<T, SD>(jb.s[1], L, P, cnt, TPT, A, B, sink);
        const T sc = SD > 0 ? T(1) / T(jb.m) : T(1);
        cx<T>* __restrict__ y = (cx<T>*)a.out + jb.out_off;
        for (int i = tid; i < cnt * L; i += RAG_THREADS) {
            const int k2 = i / cnt, c = i - k2 * cnt;
            y[(r0 + c) + (long long)m1 * k2] = scl(F[c * P + k2], sc);
        }
        return;
    }
    const cx<T>* __restrict__ filt = (const cx<T>*)jb.filt[a.dir > 0 ? 1 : 0];
    cx<T>* C = rag_fin(jb.s[1].npass, A, B);
    cx<T>* D = C == A ? B : A;
    auto sink1 = [&](int c, int k2, cx<T> v) { C[c * P + k2] = cmul(v, filt[(long long)(r0 + c) * L + k2]); };
    rag_passes<T, -1>(jb.s[1], L, P, cnt, TPT, A, B, sink1);
    cx<T>* E = rag_fin(jb.s[1].npass, C, D);
    auto sink2 = [&](int c, int p, cx<T> v) {
        E[c * P + p] = twmul<1>(v, rag_tw<T>(jb, (long long)(r0 + c) * p));
    };
    rag_passes<T, 1>(jb.s[1], L, P, cnt, TPT, C, D, sink2);
    for (int i = tid; i < cnt * L; i += RAG_THREADS) {
        const int c = i / L, p = i - c * L;
        S[(long long)(r0 + c) * L + p] = E[c * P + p];
    }
}

// Second column task of a split Bluestein job (phase 2): columns p0 ..
// p0+cnt-1 of S, inverse FFT over k1, output z[p + m2*q] times the
// post-chirp (and 1/n for the inverse transform), cropped to n.
template <typename T>
__device__ void rag_cols2(const RagArgs a, const RJob& jb, const RTask& tk, cx<T>* lds) {
    const int L = jb.m1, P = L + 1, TPT = RAG_THREADS / jb.g_col;
    const int cnt = tk.count, p0 = tk.first, m2 = jb.m2, tid = threadIdx.x;
    const cx<T>* __restrict__ S = (const cx<T>*)a.scratch + jb.scr_off;
    const cx<T>* __restrict__ chirp = (const cx<T>*)jb.chirp;
    cx<T>* __restrict__ y = (cx<T>*)a.out + jb.out_off;
    cx<T>* A = lds;
    cx<T>* B = lds + RAG_HALF;
    for (int i = tid; i < cnt * L; i += RAG_THREADS) {
        const int k1 = i / cnt, c = i - k1 * cnt;
        A[c * P + k1] = S[(long long)k1 * m2 + p0 + c];
    }
    __syncthreads();
    cx<T>* F = rag_fin(jb.s[0].npass, A, B);
    auto sink = [&](int c, int q, cx<T> v) { F[c * P + q] = v; };
    rag_passes<T, 1>(jb.s[0], L, P, cnt, TPT, A, B, sink);
    const bool inv = a.dir > 0;
    const T sc = inv ? T(1) / T(jb.n) : T(1);
    for (int i = tid; i < cnt * L; i += RAG_THREADS) {
        const int q = i / cnt, c = i - q * cnt;
        const long long n = p0 + c + (long long)m2 * q;
        if (n < jb.n) {
            const cx<T> v = F[c * P + q];
            y[n] = scl(inv ? cmulc(v, chirp[n]) : cmul(v, chirp[n]), sc);
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(RAG_THREADS) k_ragged(RagArgs a) {
    __shared__ cx<T> lds[2 * RAG_HALF];
    __shared__ int s_task;
    const int tid = threadIdx.x;
    while (true) {
        if (tid == 0) s_task = atomicAdd(a.counter, 1);
        __syncthreads();
        const int id = s_task;
        __syncthreads();
        if (id >= a.ntasks) break;
        const RTask tk = a.tasks[id];
        const RJob& jb = a.jobs[tk.job];
        const bool inv = a.dir > 0;
        switch (tk.type) {
        case RT_WHOLE:
            rag_whole<T>(a, jb, lds);
            break;
        case RT_COLS:
            if (inv && !jb.blue) rag_cols<T, 1>(a, jb, tk, lds);
            else rag_cols<T, -1>(a, jb, tk, lds);
            break;
        case RT_ROWS:
            if (inv && !jb.blue) rag_rows<T, 1>(a, jb, tk, lds);
            else rag_rows<T, -1>(a, jb, tk, lds);
            break;
        default:
            rag_cols2<T>(a, jb, tk, lds);
            break;
        }
    }
}

// Clears the phase counters (a kernel, so the launches can be captured in a
// graph and replayed).
__global__ void k_rag_reset(int* counter) {
    if (threadIdx.x < 4) counter[threadIdx.x] = 0;
}

}  // namespace afft
