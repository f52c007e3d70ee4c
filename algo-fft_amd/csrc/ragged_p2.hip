// Instantiations of the ragged batch's power-of-two task kernels (one per
// task type, sub-transform length, direction and precision), in their own
// translation unit so they compile beside the generic ones.
#include "kragged.hpp"

namespace afft {

template <typename T, int L>
static const void* p2_kernel_l(int kind, int dir) {
    if constexpr (L >= RAG_P2_COL_MIN && L <= RAG_P2_COL_MAX) {
        switch (kind) {
        case RL_COLS_P2:
            return dir > 0 ? (const void*)&k_rag_cols_p2<T, L, 1, false> : (const void*)&k_rag_cols_p2<T, L, -1, false>;
        case RL_COLS_BLUE: return (const void*)&k_rag_cols_p2<T, L, -1, true>;
        case RL_COLS2: return (const void*)&k_rag_cols2<T, L>;
        default: break;
        }
    }
    if constexpr (L >= RAG_P2_ROW_MIN && L <= RAG_P2_ROW_MAX) {
        switch (kind) {
        case RL_ROWS_P2: return dir > 0 ? (const void*)&k_rag_rows_p2<T, L, 1> : (const void*)&k_rag_rows_p2<T, L, -1>;
        case RL_ROWS_BLUE: return (const void*)&k_rag_rows_blue<T, L>;
        default: break;
        }
    }
    return nullptr;
}

template <typename T>
static const void* p2_kernel_t(int kind, int L, int dir) {
    switch (L) {
    case 16: return p2_kernel_l<T, 16>(kind, dir);
    case 32: return p2_kernel_l<T, 32>(kind, dir);
    case 64: return p2_kernel_l<T, 64>(kind, dir);
    case 128: return p2_kernel_l<T, 128>(kind, dir);
    case 256: return p2_kernel_l<T, 256>(kind, dir);
    case 512: return p2_kernel_l<T, 512>(kind, dir);
    case 1024: return p2_kernel_l<T, 1024>(kind, dir);
    case 2048: return p2_kernel_l<T, 2048>(kind, dir);
    case 4096: return p2_kernel_l<T, 4096>(kind, dir);
    default: return nullptr;
    }
}

// The kernel of a power-of-two split launch (RL_COLS_P2, RL_COLS_BLUE,
// RL_ROWS_P2, RL_ROWS_BLUE, RL_COLS2) for sub-transforms of length L, or
// nullptr.
const void* rag_p2_kernel(int prec, int kind, int L, int dir) {
    return prec ? p2_kernel_t<double>(kind, L, dir) : p2_kernel_t<float>(kind, L, dir);
}

}  // namespace afft
