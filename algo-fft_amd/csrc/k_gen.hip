// Runtime-size mixed-radix kernels and the strided copy kernel.
#include "kgen.hpp"

namespace afft {

const void* gen_kernel(int prec, int var, int dir) {
#define AFFT_GEN_CASE(T, P)                                                             \
    if (prec == P) {                                                                    \
        switch (var * 2 + (dir > 0)) {                                                  \
        case 0: return (const void*)&k_gen<T, VAR_ROW, -1>;                             \
        case 1: return (const void*)&k_gen<T, VAR_ROW, 1>;                              \
        case 2: return (const void*)&k_gen<T, VAR_COL, -1>;                             \
        case 3: return (const void*)&k_gen<T, VAR_COL, 1>;                              \
        case 4: return (const void*)&k_gen<T, VAR_4A, -1>;                              \
        case 5: return (const void*)&k_gen<T, VAR_4A, 1>;                               \
        case 6: return (const void*)&k_gen<T, VAR_4B, -1>;                              \
        case 7: return (const void*)&k_gen<T, VAR_4B, 1>;                               \
        }                                                                               \
    }
    AFFT_GEN_CASE(float, 0)
    AFFT_GEN_CASE(double, 1)
#undef AFFT_GEN_CASE
    return nullptr;
}

const void* gen_ip_kernel(int prec, int dir) {
    if (prec == 0) return dir < 0 ? (const void*)&k_gen_ip<float, -1> : (const void*)&k_gen_ip<float, 1>;
    return dir < 0 ? (const void*)&k_gen_ip<double, -1> : (const void*)&k_gen_ip<double, 1>;
}

// VAR_COL, VAR_4A or VAR_4B only.
const void* gen_ipg_kernel(int prec, int var, int dir) {
    if (var == VAR_COL) {
        if (prec == 0) return dir < 0 ? (const void*)&k_gen_ipg<float, VAR_COL, -1> : (const void*)&k_gen_ipg<float, VAR_COL, 1>;
        return dir < 0 ? (const void*)&k_gen_ipg<double, VAR_COL, -1> : (const void*)&k_gen_ipg<double, VAR_COL, 1>;
    }
    if (var != VAR_4A && var != VAR_4B) return nullptr;
    if (prec == 0) {
        if (var == VAR_4A) return dir < 0 ? (const void*)&k_gen_ipg<float, VAR_4A, -1> : (const void*)&k_gen_ipg<float, VAR_4A, 1>;
        return dir < 0 ? (const void*)&k_gen_ipg<float, VAR_4B, -1> : (const void*)&k_gen_ipg<float, VAR_4B, 1>;
    }
    if (var == VAR_4A) return dir < 0 ? (const void*)&k_gen_ipg<double, VAR_4A, -1> : (const void*)&k_gen_ipg<double, VAR_4A, 1>;
    return dir < 0 ? (const void*)&k_gen_ipg<double, VAR_4B, -1> : (const void*)&k_gen_ipg<double, VAR_4B, 1>;
}

const void* copy_kernel(int prec) {
    return prec == 0 ? (const void*)&k_copy<float> : (const void*)&k_copy<double>;
}

const void* rpost_kernel(int prec, int dir) {
    if (prec == 0) return dir < 0 ? (const void*)&k_rpost<float, -1> : (const void*)&k_rpost<float, 1>;
    return dir < 0 ? (const void*)&k_rpost<double, -1> : (const void*)&k_rpost<double, 1>;
}

}  // namespace afft
