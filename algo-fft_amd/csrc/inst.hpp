// Kernel instantiation lists.  Each translation unit includes this header and
// expands AFFT_POW2(T, VAR, MAXN) for one precision and variant; the
// registrars run at library load and fill the kernel registry.
#pragma once
#include "kernels.hpp"

namespace afft {

// Transforms per block: row blocks aim at 256 threads; column blocks take up
// to 32 adjacent columns (256-byte rows of complex64), the four-step passes
// up to 16, within 1024 threads and 128 KiB of LDS.  Measured against 16
// columns: float32 R2C 512^3 0.841 -> 0.747 ms (its middle axis has rows 257
// points wide, so 128-byte column groups straddle cache lines), complex128
// 8192^2 1.295 -> 1.276 ms; the four-step passes lose about 1 % with 32.
template <typename T, int N, int E, int VAR>
constexpr int tpb_for() {
    constexpr int TT = N / E;
    if constexpr (VAR == VAR_ROW || VAR == VAR_BLUE || VAR == VAR_REAL) {
        return TT >= 256 ? 1 : 256 / TT;
    } else {
        int t = VAR == VAR_COL ? 32 : 16;
        while (t > 1 && (t * TT > 1024 || t * N * (int)sizeof(cx<T>) > 131072)) t /= 2;
        return t;
    }
}

}  // namespace afft

#define AFFT_CAT2(a, b) a##b
#define AFFT_CAT(a, b) AFFT_CAT2(a, b)

// One (schedule, variant) pair in both directions.
#define AFFT_REG(T, VAR, E, ...)                                                                      \
    static ::afft::Registrar<T, ::afft::Sched<__VA_ARGS__>, E,                                        \
                             ::afft::tpb_for<T, ::afft::Sched<__VA_ARGS__>::N, E, VAR>(), VAR, -1>    \
        AFFT_CAT(reg_f_, __COUNTER__);                                                                \
    static ::afft::Registrar<T, ::afft::Sched<__VA_ARGS__>, E,                                        \
                             ::afft::tpb_for<T, ::afft::Sched<__VA_ARGS__>::N, E, VAR>(), VAR, 1>     \
        AFFT_CAT(reg_i_, __COUNTER__);

// Power-of-two schedules: radix-16 passes with the remainder radix first
// (the first pass needs no twiddles).
#define AFFT_POW2_SMALL(T, VAR)                 \
    AFFT_REG(T, VAR, 2, 2)                      \
    AFFT_REG(T, VAR, 4, 4)                      \
    AFFT_REG(T, VAR, 8, 8)                      \
    AFFT_REG(T, VAR, 16, 16)                    \
    AFFT_REG(T, VAR, 16, 2, 16)                 \
    AFFT_REG(T, VAR, 16, 4, 16)                 \
    AFFT_REG(T, VAR, 16, 8, 16)                 \
    AFFT_REG(T, VAR, 16, 16, 16)                \
    AFFT_REG(T, VAR, 16, 2, 16, 16)             \
    AFFT_REG(T, VAR, 16, 4, 16, 16)             \
    AFFT_REG(T, VAR, 16, 8, 16, 16)             \
    AFFT_REG(T, VAR, 16, 16, 16, 16)            \
    AFFT_REG(T, VAR, 16, 2, 16, 16, 16)

#define AFFT_POW2_16K(T, VAR) AFFT_REG(T, VAR, 16, 4, 16, 16, 16)
