// Persistent work-queue kernel for a ragged batch: thousands of independent
// 1-D complex transforms of different lengths (powers of two, 2/3/5/7-smooth
// and Bluestein sizes) in one launch.
//
// Each block takes the next job from a global counter (jobs are sorted by
// decreasing cost on the host, so the queue is a longest-processing-time
// schedule) and runs the whole transform with the runtime Stockham passes of
// kgen.hpp.  Jobs whose FFT length m fits in LDS (2 * m <= RAG_LDS points)
// ping-pong between two LDS buffers; larger ones ping-pong between two
// per-block scratch buffers in global memory (L2-resident at these sizes).
// Bluestein jobs run the chain pre-chirp -> FFT_m -> filter -> IFFT_m ->
// post-chirp inside the block.  The loop ends when the counter passes the
// job count, which every block reaches.
#pragma once
#include "kgen.hpp"

namespace afft {

constexpr int RAG_THREADS = 256;
constexpr int RAG_LDS = 8192;       // complex points of LDS per block (two buffers of 4096)
constexpr int RAG_MAXPASS = 16;

struct RJob {
    long long in_off, out_off;   // element offsets of the job's input and output
    int n;                       // logical length
    int m;                       // FFT length (n, or the Bluestein padding)
    int blue;                    // 1: Bluestein chain
    int npass;
    int rad[RAG_MAXPASS];
    int ns[RAG_MAXPASS];
    int twoff[RAG_MAXPASS];
    const void* tw;              // schedule twiddles of m
    const void* chirp;           // Bluestein: n entries
    const void* filt[2];         // Bluestein: m entries each, [forward, inverse], 1/m folded in
};

struct RagArgs {
    const void* in;
    void* out;
    const RJob* jobs;
    int njobs;
    int dir;
    int* counter;
    void* scratch;               // 2 * scratch_pts complex per block
    long long scratch_pts;
};

// One Stockham transform of length m over runtime passes.
//   src(n)  : input point n (first pass)
//   sink(k, y): output point k of the last pass
// Intermediate passes ping-pong between bufA and bufB (LDS or global).
template <typename T, int DIR, class Src, class Sink>
__device__ __forceinline__ void rag_fft(const RJob& jb, int m, cx<T>* bufA, cx<T>* bufB, Src src, Sink sink) {
    const int tid = threadIdx.x;
    const cx<T>* __restrict__ tw = (const cx<T>*)jb.tw;
    for (int p = 0; p < jb.npass; ++p) {
        const int R = jb.rad[p];
        const int NS = jb.ns[p];
        const bool last = p == jb.npass - 1;
        cx<T>* rd = (p & 1) ? bufA : bufB;   // written by pass p-1
        cx<T>* wr = (p & 1) ? bufB : bufA;
        auto load = [&](int n) -> cx<T> { return p == 0 ? src(n) : rd[n]; };
        auto store = [&](int k, cx<T> v) {
            if (last) sink(k, v);
            else wr[k] = v;
        };
        if (R == 1) {
            for (int j = tid; j < m; j += RAG_THREADS) store(j, load(j));
        } else {
            rt_pass_any<DIR>(R, tid, RAG_THREADS, m, NS, tw + jb.twoff[p], load, store);
        }
        __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(RAG_THREADS) k_ragged(RagArgs a) {
    __shared__ cx<T> lds[RAG_LDS];
    __shared__ int s_job;
    const int tid = threadIdx.x;
    const cx<T>* __restrict__ gin = (const cx<T>*)a.in;
    cx<T>* __restrict__ gout = (cx<T>*)a.out;
    cx<T>* scr = (cx<T>*)a.scratch + (long long)blockIdx.x * 2 * a.scratch_pts;
    const int di = a.dir > 0 ? 1 : 0;
    while (true) {
        if (tid == 0) s_job = atomicAdd(a.counter, 1);
        __syncthreads();
        const int jid = s_job;
        __syncthreads();
        if (jid >= a.njobs) break;
        const RJob& jb = a.jobs[jid];
        const cx<T>* x = gin + jb.in_off;
        cx<T>* y = gout + jb.out_off;
        const int m = jb.m;
        const bool in_lds = 2 * m <= RAG_LDS;
        cx<T>* A = in_lds ? lds : scr;
        cx<T>* B = in_lds ? lds + m : scr + a.scratch_pts;
        if (!jb.blue) {
            const T sc = a.dir > 0 ? T(1) / T(m) : T(1);
            auto src = [&](int n) { return x[n]; };
            auto sink = [&](int k, cx<T> v) { y[k] = a.dir > 0 ? scl(v, sc) : v; };
            if (a.dir > 0) rag_fft<T, 1>(jb, m, A, B, src, sink);
            else rag_fft<T, -1>(jb, m, A, B, src, sink);
        } else {
            // pre-chirp and zero padding on the load, filter product on the
            // forward FFT's last store, post-chirp on the inverse FFT's
            const cx<T>* chirp = (const cx<T>*)jb.chirp;
            const cx<T>* filt = (const cx<T>*)jb.filt[di];
            const int n = jb.n;
            const bool inv = a.dir > 0;
            const T sc = inv ? T(1) / T(n) : T(1);
            // The forward FFT's last pass reads buffer ((P-1)&1 ? A : B), so its
            // filtered output goes to the other one, C.  The inverse FFT's
            // first pass reads C and writes D (the buffer that is not C); C is
            // free again from its second pass on.
            const int P = jb.npass;
            cx<T>* C = ((P - 1) & 1) ? B : A;
            cx<T>* D = C == A ? B : A;
            auto src1 = [&](int k) {
                if (k >= n) return mk<T>(0, 0);
                return inv ? cmulc(x[k], chirp[k]) : cmul(x[k], chirp[k]);
            };
            auto sink1 = [&](int k, cx<T> v) { C[k] = cmul(v, filt[k]); };
            rag_fft<T, -1>(jb, m, A, B, src1, sink1);
            auto src2 = [&](int k) { return C[k]; };
            auto sink2 = [&](int k, cx<T> v) {
                if (k < n) y[k] = scl(inv ? cmulc(v, chirp[k]) : cmul(v, chirp[k]), sc);
            };
            rag_fft<T, 1>(jb, m, D, C, src2, sink2);
        }
    }
}

// Clears the queue counter (a kernel, so the launch pair can be captured in
// a graph and replayed).
__global__ void k_rag_reset(int* counter) {
    if (threadIdx.x == 0) counter[0] = 0;
}

}  // namespace afft
