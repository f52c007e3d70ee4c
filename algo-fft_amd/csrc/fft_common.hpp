// Complex arithmetic and small in-register DFT codelets for gfx950.
//
// Every codelet computes an in-place, natural-order DFT of R points held in
// registers.  DIR = -1 is the forward transform (exp(-2πi nk/R), the sign of
// algo-fft's Plan.Forward, plan.go:325), DIR = +1 the unnormalised inverse.
// Constants are evaluated at compile time in double precision (constexpr
// Taylor series on a reduced argument) and rounded once to the working type.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

namespace afft {

template <typename T>
struct __attribute__((aligned(2 * sizeof(T)))) cx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> mk(T a, T b) { cx<T> r; r.x = a; r.y = b; return r; }

// A read-only pointer into global memory.  A pointer that a kernel reads out
// of a table in memory, rather than receives as an argument, is generic to
// the compiler, which then reads through it with flat loads: those count
// against the LDS operations as well, so every wait for one also drains the
// LDS reads and writes in flight.  Reading through a gptr instead makes them
// global loads.  It indexes and offsets like a const cx<U>*, and an element
// read returns a value (copying the address-space-qualified object itself
// is not allowed in host compilation).
template <typename T>
struct gptr;
template <typename U>
struct gptr<cx<U>> {
    const __attribute__((address_space(1))) cx<U>* p;
    __device__ __forceinline__ cx<U> operator[](long long i) const {
        cx<U> r;
        r.x = p[i].x;
        r.y = p[i].y;
        return r;
    }
    __device__ __forceinline__ gptr operator+(long long i) const { return gptr{p + i}; }
};
template <typename T>
__device__ __forceinline__ gptr<T> as_global(const void* p) {
    return gptr<T>{(const __attribute__((address_space(1))) T*)p};
}

// Non-temporal flags of a launch (KArgs::nt): streaming stores of the
// output, and streaming loads of the input, for data too large to be re-read
// from the last-level cache.
constexpr int NT_STORE = 1, NT_LOAD = 2;

// Output store; non-temporal under NT_STORE.
template <typename T>
__device__ __forceinline__ void st_out(cx<T>* p, cx<T> v, int nt) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    if (nt & NT_STORE) {
        const v2 val = {v.x, v.y};
        __builtin_nontemporal_store(val, (v2*)p);
    } else {
        *p = v;
    }
}
// j mod d for 0 <= j < 2^31 and d >= 1, given m = mod_magic(d) =
// floor((2^32 - 1) / d).  q = mulhi(j, m) satisfies
// j/d - j(1+d)/(d 2^32) < q <= j/d, and j(1+d)/(d 2^32) < 1, so q is
// floor(j/d) or one less and a single correction yields the remainder.
// Replaces the ~20-instruction integer division of a runtime modulus.
__host__ __device__ constexpr unsigned mod_magic(unsigned d) { return 0xFFFFFFFFu / d; }
__device__ __forceinline__ int fast_mod(int j, int d, unsigned m) {
    const unsigned q = __umulhi((unsigned)j, m);
    const int r = j - (int)(q * (unsigned)d);
    return r >= d ? r - d : r;
}

// Non-temporal input load.
template <typename T>
__device__ __forceinline__ cx<T> ld_nt(const cx<T>* p) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 d = __builtin_nontemporal_load((const v2*)p);
    return mk<T>(d.x, d.y);
}

template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cx<T> cmulc(cx<T> a, cx<T> b) {
    return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// a * w for DIR=-1, a * conj(w) for DIR=+1: tables hold forward twiddles.
template <int DIR, typename T>
__device__ __forceinline__ cx<T> twmul(cx<T> a, cx<T> w) {
    if constexpr (DIR < 0) return cmul(a, w); else return cmulc(a, w);
}
template <typename T>
__device__ __forceinline__ cx<T> scl(cx<T> a, T s) { return mk(a.x * s, a.y * s); }
template <typename T>
__device__ __forceinline__ cx<T> conjx(cx<T> a) { return mk(a.x, -a.y); }
// multiply by DIR*i: -i for DIR=-1, +i for DIR=+1
template <int DIR, typename T>
__device__ __forceinline__ cx<T> rot(cx<T> a) {
    if constexpr (DIR < 0) return mk(a.y, -a.x); else return mk(-a.y, a.x);
}

// ---- compile-time trigonometry --------------------------------------------
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double taylor_sin(double x) {
    double term = x, sum = x, x2 = x * x;
    for (int i = 1; i < 14; ++i) {
        term *= -x2 / double((2 * i) * (2 * i + 1));
        sum += term;
    }
    return sum;
}
constexpr double taylor_cos(double x) {
    double term = 1.0, sum = 1.0, x2 = x * x;
    for (int i = 1; i < 14; ++i) {
        term *= -x2 / double((2 * i - 1) * (2 * i));
        sum += term;
    }
    return sum;
}
// cos(2π m / r) and sin(2π m / r) for integers, reduced to the first octant
// so the series stays accurate to the last bit of a double.
constexpr double cos2pi(long m, long r) {
    m %= r; if (m < 0) m += r;
    // use symmetry: angle = 2π m / r in [0, 2π)
    long m8 = 8 * m;               // octant index = floor(8m / r)
    if (m8 <= r) return taylor_cos(2.0 * kPi * double(m) / double(r));
    if (m8 <= 2 * r) return taylor_sin(2.0 * kPi * double(r - 4 * m) / double(4 * r));
    if (m8 <= 4 * r) return -cos2pi(r - 2 * m, 2 * r);
    return cos2pi(r - m, r);
}
constexpr double sin2pi(long m, long r) {
    m %= r; if (m < 0) m += r;
    long m8 = 8 * m;
    if (m8 <= r) return taylor_sin(2.0 * kPi * double(m) / double(r));
    if (m8 <= 2 * r) return taylor_cos(2.0 * kPi * double(r - 4 * m) / double(4 * r));
    if (m8 <= 4 * r) return sin2pi(r - 2 * m, 2 * r);
    return -sin2pi(r - m, r);
}

// v * W_R^{M} with W_R = exp(DIR * 2πi / R); M is a compile-time constant,
// so trivial factors (±1, ±i, (±1±i)/√2) cost no multiplies.
template <int R, int M, int DIR, typename T>
__device__ __forceinline__ cx<T> mulw(cx<T> v) {
    constexpr int m = ((M % R) + R) % R;
    if constexpr (m == 0) {
        return v;
    } else if constexpr (2 * m == R) {
        return mk(-v.x, -v.y);
    } else if constexpr (4 * m == R) {
        return rot<DIR>(v);                   // W_R^{R/4} = exp(DIR*iπ/2) = DIR*i
    } else if constexpr (4 * m == 3 * R) {
        return rot<-DIR>(v);                  // W_R^{3R/4} = -DIR*i
    } else if constexpr (8 * m == R || 8 * m == 3 * R || 8 * m == 5 * R || 8 * m == 7 * R) {
        constexpr T h = T(0.70710678118654752440084436210484904);
        constexpr double c = cos2pi(m, R);
        constexpr double s = DIR * sin2pi(m, R);
        // (c + i s) with |c| = |s| = √½
        constexpr T cs = c > 0 ? h : -h;
        constexpr T ss = s > 0 ? h : -h;
        return mk(cs * v.x - ss * v.y, cs * v.y + ss * v.x);
    } else {
        constexpr T c = T(cos2pi(m, R));
        constexpr T s = T(DIR * sin2pi(m, R));
        return mk(c * v.x - s * v.y, c * v.y + s * v.x);
    }
}

// ---- small DFT codelets ------------------------------------------------------
// dft<R, DIR>(v, S): in-place DFT of v[0], v[S], ..., v[(R-1)S].

template <int R, int DIR, int S = 1, typename T>
__device__ __forceinline__ void dft(cx<T>* v);

template <int DIR, int S, typename T>
__device__ __forceinline__ void dft2(cx<T>* v) {
    cx<T> a = v[0], b = v[S];
    v[0] = a + b;
    v[S] = a - b;
}

template <int DIR, int S, typename T>
__device__ __forceinline__ void dft4(cx<T>* v) {
    cx<T> t0 = v[0] + v[2 * S], t1 = v[0] - v[2 * S];
    cx<T> t2 = v[S] + v[3 * S], t3 = rot<DIR>(v[S] - v[3 * S]);
    // rot<DIR> multiplies by -i for the forward transform
    v[0] = t0 + t2;
    v[2 * S] = t0 - t2;
    v[S] = t1 + t3;
    v[3 * S] = t1 - t3;
}

// Odd prime radix by the conjugate-pair form:
//   y_k = a0 + Σ_j c_jk b_j  +  DIR*i Σ_j s_jk d_j,   b_j = a_j + a_{R-j}, d_j = a_j - a_{R-j}
// cos / sin of 2*pi*i/R for i < R, built at compile time
template <int R>
struct RootTable {
    double c[R], s[R];
    constexpr RootTable() : c(), s() {
        for (int i = 0; i < R; ++i) {
            c[i] = cos2pi(i, R);
            s[i] = sin2pi(i, R);
        }
    }
};

template <int R, int DIR, int S, typename T>
__device__ __forceinline__ void dft_odd(cx<T>* v) {
    constexpr int H = (R - 1) / 2;
    constexpr RootTable<R> tab{};
    cx<T> b[H], d[H];
    cx<T> a0 = v[0];
    cx<T> y0 = a0;
#pragma unroll
    for (int j = 1; j <= H; ++j) {
        b[j - 1] = v[j * S] + v[(R - j) * S];
        d[j - 1] = v[j * S] - v[(R - j) * S];
        y0 = y0 + b[j - 1];
    }
#pragma unroll
    for (int k = 1; k <= H; ++k) {
        T rx = a0.x, ry = a0.y, ix = 0, iy = 0;
#pragma unroll
        for (int j = 1; j <= H; ++j) {
            const T c = T(tab.c[(j * k) % R]);
            const T s = T(tab.s[(j * k) % R]);
            rx += c * b[j - 1].x;
            ry += c * b[j - 1].y;
            ix += s * d[j - 1].x;
            iy += s * d[j - 1].y;
        }
        // forward (DIR=-1): y_k = r - i*(ix,iy); y_{R-k} = r + i*(ix,iy)
        // i*(ix + i iy) = (-iy, ix)
        if constexpr (DIR < 0) {
            v[k * S] = mk(rx + iy, ry - ix);
            v[(R - k) * S] = mk(rx - iy, ry + ix);
        } else {
            v[k * S] = mk(rx - iy, ry + ix);
            v[(R - k) * S] = mk(rx + iy, ry - ix);
        }
    }
    v[0] = y0;
}

// Twiddle step of the composite codelet: w[k1*R2 + n2] *= W_R^{k1*n2}, with
// every exponent a compile-time constant.
template <int R, int R2, int DIR, int K1, typename T, int... N2>
__device__ __forceinline__ void tw_row(cx<T>* w, std::integer_sequence<int, N2...>) {
    ((w[K1 * R2 + N2] = mulw<R, K1 * N2, DIR>(w[K1 * R2 + N2])), ...);
}
template <int R, int R2, int DIR, typename T, int... K1>
__device__ __forceinline__ void tw_all(cx<T>* w, std::integer_sequence<int, K1...>) {
    (tw_row<R, R2, DIR, K1>(w, std::make_integer_sequence<int, R2>{}), ...);
}

// Composite radix R = R1 * R2 with input index n = n1*R2 + n2 and output
// index k = k1 + R1*k2:  inner DFTs of size R1 over n1, twiddle W_R^{n2 k1},
// outer DFTs of size R2 over n2.  Works on a strided view; result is written
// back in natural order through a register transpose.
template <int R1, int R2, int DIR, int S, typename T>
__device__ __forceinline__ void dft_composite(cx<T>* v) {
    constexpr int R = R1 * R2;
    cx<T> w[R];
#pragma unroll
    for (int i = 0; i < R; ++i) w[i] = v[i * S];
    // inner: for each n2, DFT over n1 (stride R2 in w)
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) dft<R1, DIR, R2>(w + n2);
    // now w[k1*R2 + n2] = Y[k1, n2]; twiddle
    tw_all<R, R2, DIR>(w, std::make_integer_sequence<int, R1>{});
    // outer: for each k1, DFT over n2 (contiguous in w)
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) dft<R2, DIR, 1>(w + k1 * R2);
    // w[k1*R2 + k2] = X[k1 + R1*k2]
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1)
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) v[(k1 + R1 * k2) * S] = w[k1 * R2 + k2];
}

template <int R, int DIR, int S, typename T>
__device__ __forceinline__ void dft(cx<T>* v) {
    if constexpr (R == 1) {
    } else if constexpr (R == 2) {
        dft2<DIR, S>(v);
    } else if constexpr (R == 4) {
        dft4<DIR, S>(v);
    } else if constexpr (R == 3 || R == 5 || R == 7 || R == 11 || R == 13) {
        dft_odd<R, DIR, S>(v);
    } else if constexpr (R == 8) {
        dft_composite<2, 4, DIR, S>(v);
    } else if constexpr (R == 16) {
        dft_composite<4, 4, DIR, S>(v);
    } else if constexpr (R == 32) {
        dft_composite<4, 8, DIR, S>(v);
    } else if constexpr (R == 6) {
        dft_composite<2, 3, DIR, S>(v);
    } else if constexpr (R == 9) {
        dft_composite<3, 3, DIR, S>(v);
    } else if constexpr (R == 10) {
        dft_composite<2, 5, DIR, S>(v);
    } else if constexpr (R == 12) {
        dft_composite<4, 3, DIR, S>(v);
    } else if constexpr (R == 15) {
        dft_composite<3, 5, DIR, S>(v);
    } else {
        static_assert(R == 0, "unsupported radix");
    }
}

}  // namespace afft
