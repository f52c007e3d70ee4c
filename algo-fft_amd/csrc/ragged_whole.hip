// Instantiations of the ragged batch's whole-job kernels (per length on the
// compile-time cores from 2048 points, one runtime-pass kernel for the other
// lengths, and the Bluestein forms of both), in their own translation unit
// so they compile beside the split kernels.
#include "kragged.hpp"

namespace afft {

constexpr int RAG_WHOLE_P2_MIN = 2048;   // shortest length on the compile-time cores

template <typename T, int L>
static const void* whole_p2_l(int kind, int dir, int* threads) {
    if constexpr (L * (int)sizeof(cx<T>) <= 16384 * 8) {   // the whole transform in 128 KiB of LDS
        if (kind == RL_WHOLE_P2) {
            *threads = P2Whole<T, L>::THREADS;
            return dir > 0 ? (const void*)&k_rag_whole_p2<T, L, 1> : (const void*)&k_rag_whole_p2<T, L, -1>;
        }
        if (kind == RL_WHOLE_BLUE) {
            *threads = P2WholeBlue<T, L>::THREADS;
            return (const void*)&k_rag_whole_blue_p2<T, L>;
        }
    }
    return nullptr;
}

template <typename T>
static const void* whole_kernel_t(int kind, int L, int dir, int* threads) {
    *threads = RAG_THREADS;
    if (L == 0) {
        switch (kind) {
        case RL_WHOLE:
            *threads = RagW<T>::THREADS;
            return dir > 0 ? (const void*)&k_rag_whole<T, 1> : (const void*)&k_rag_whole<T, -1>;
        case RL_WHOLE_BLUE: return (const void*)&k_rag_whole_blue<T>;
        default: return nullptr;
        }
    }
    switch (L) {
    case 2048: return whole_p2_l<T, 2048>(kind, dir, threads);
    case 4096: return whole_p2_l<T, 4096>(kind, dir, threads);
    case 8192: return whole_p2_l<T, 8192>(kind, dir, threads);
    case 16384: return whole_p2_l<T, 16384>(kind, dir, threads);
    default: return nullptr;
    }
}

// The kernel of a whole-job launch and its threads per block: RL_WHOLE
// (L = 0, every length on the runtime passes), RL_WHOLE_BLUE with L = 0
// (paddings below RAG_WHOLE_P2_MIN on the runtime passes, one job per
// block), or RL_WHOLE_P2 / RL_WHOLE_BLUE of one length L >= RAG_WHOLE_P2_MIN
// on the compile-time cores; nullptr if there is none.
const void* rag_whole_kernel(int prec, int kind, int L, int dir, int* threads) {
    return prec ? whole_kernel_t<double>(kind, L, dir, threads) : whole_kernel_t<float>(kind, L, dir, threads);
}

int rag_whole_p2_min() { return RAG_WHOLE_P2_MIN; }

}  // namespace afft
