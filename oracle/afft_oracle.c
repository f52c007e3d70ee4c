/*
 * algo-fft oracle — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C CPU restatement of the reference library's FFT algorithms
 * (github.com/cwbudde/algo-fft, pure Go).  It is used in exactly three
 * places: the CPU-side tests (as the parity checker), the smoke check, and
 * the CPU-baseline leg of the benchmark.  The HIP library never links,
 * loads or calls it.
 *
 * Each function cites the reference file:line it follows.  The same body
 * (fft_impl.inc) is compiled twice: in float for the reference's complex64
 * plans and in double for its complex128 plans.
 *
 * Deliberate difference from the reference, reported rather than matched:
 * the Bluestein chirp exp(-iπk²/n) defaults to an integer reduction of k²
 * mod 2n before the FP64 cos/sin (exact_chirp=1); exact_chirp=0 reproduces
 * the reference's unreduced formula (internal/kernels/bluestein.go:23-28).
 *
 * Pinning: see the tests under tests/ (closed forms, long-double brute
 * force DFT, numpy.fft golden vectors, the reference's own known-answer
 * tests for twiddles, trivial DFTs and convolution).
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ---- portable seeded generator ---------------------------------------
 * Counter-based splitmix64: value i of stream `seed` is
 *   z = splitmix64(seed * 0xD1342543DE82EF95 + i)
 *   u = (z >> 11) * 2^-53           in [0, 1)
 * The same stream is produced by algofft_hip.synth (numpy) so CPU and GPU
 * tests draw identical inputs. */
static inline uint64_t or_splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void or_rng_u01(uint64_t seed, uint64_t offset, long count, double *out) {
    uint64_t base = seed * 0xD1342543DE82EF95ULL + offset;
    for (long i = 0; i < count; i++)
        out[i] = (double)(or_splitmix64(base + (uint64_t)i) >> 11) * (1.0 / 9007199254740992.0);
}

/* internal/reference/dft.go:16-51 (NaiveDFT) and its inverse :53-100,
 * evaluated in long double with the exact index reduction (k*j mod n).
 * inverse=1 applies the 1/n scale like NaiveIDFT. */
int or_dft_naive(long n, const double *in, double *out, int inverse) {
    if (n < 1) return -1;
    long double *c = (long double *)malloc(sizeof(long double) * (size_t)n);
    long double *s = (long double *)malloc(sizeof(long double) * (size_t)n);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (long k = 0; k < n; k++) {
        long double a = (inverse ? 2.0L : -2.0L) * pi * (long double)k / (long double)n;
        c[k] = cosl(a);
        s[k] = sinl(a);
    }
    for (long k = 0; k < n; k++) {
        long double sr = 0, si = 0;
        long idx = 0;
        for (long j = 0; j < n; j++) {
            long double xr = in[2 * j], xi = in[2 * j + 1];
            sr += xr * c[idx] - xi * s[idx];
            si += xr * s[idx] + xi * c[idx];
            idx += k;
            if (idx >= n) idx -= n;
        }
        if (inverse) { sr /= (long double)n; si /= (long double)n; }
        out[2 * k] = (double)sr;
        out[2 * k + 1] = (double)si;
    }
    free(c);
    free(s);
    return 0;
}

/* Planted-mistake switch for the tests that show the pinning checks catch
 * plausible errors (0 = none):
 *   1  forward twiddle sign flipped in the Stockham stages
 *   2  last term of every mixed-radix butterfly dropped
 *   3  Bluestein chirp index off by one
 *   4  N-D axes swapped (transposed operand)
 * Only the tests set it. */
int or_fault = 0;
void or_set_fault(int f) { or_fault = f; }

/* internal/math/factor.go:31-43 */
int or_is_highly_composite(long n) {
    if (n <= 0) return 0;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    while (n % 5 == 0) n /= 5;
    return n == 1;
}

/* internal/math/power.go:12-33 */
long or_next_pow2(long n) {
    long m = 1;
    while (m < n) m <<= 1;
    return m;
}

#define R float
#define SFX(x) x##_c64
#include "fft_impl.inc"
#undef R
#undef SFX

#define R double
#define SFX(x) x##_c128
#include "fft_impl.inc"
#undef R
#undef SFX

/* ---- cost-model restatement ---------------------------------------------
 * FP64 forward and backward pass of the 32 -> 64 -> 64 -> 1 ReLU MLP the
 * library trains on the GPU (same parameter layout: W1[32][64], b1[64],
 * W2[64][64], b2[64], W3[64], b3; loss = mean squared error).  It has no
 * counterpart in the reference; it is the checker for the GPU gradients.
 * X and T are the standardised inputs and targets. */
int or_mlp_loss_grad(const double *P, const double *X, const double *T, long n, double *loss, double *G) {
    enum { I = 32, H = 64 };
    const long W1 = 0, B1 = I * H, W2 = B1 + H, B2 = W2 + H * H, W3 = B2 + H, B3 = W3 + H, NP = B3 + 1;
    for (long p = 0; p < NP; p++) G[p] = 0;
    double L = 0;
    for (long s = 0; s < n; s++) {
        const double *x = X + s * I;
        double h1[H], h2[H], d2[H], d1[H];
        for (int j = 0; j < H; j++) {
            double a = P[B1 + j];
            for (int i = 0; i < I; i++) a += x[i] * P[W1 + i * H + j];
            h1[j] = a > 0 ? a : 0;
        }
        double y = P[B3];
        for (int j = 0; j < H; j++) {
            double a = P[B2 + j];
            for (int i = 0; i < H; i++) a += h1[i] * P[W2 + i * H + j];
            h2[j] = a > 0 ? a : 0;
            y += h2[j] * P[W3 + j];
        }
        L += (y - T[s]) * (y - T[s]);
        double dy = 2.0 * (y - T[s]) / (double)n;
        G[B3] += dy;
        for (int j = 0; j < H; j++) {
            G[W3 + j] += h2[j] * dy;
            d2[j] = h2[j] > 0 ? dy * P[W3 + j] : 0;
            G[B2 + j] += d2[j];
        }
        for (int i = 0; i < H; i++) {
            double a = 0;
            for (int j = 0; j < H; j++) {
                G[W2 + i * H + j] += h1[i] * d2[j];
                a += P[W2 + i * H + j] * d2[j];
            }
            d1[i] = h1[i] > 0 ? a : 0;
            G[B1 + i] += d1[i];
        }
        for (int k = 0; k < I; k++)
            for (int i = 0; i < H; i++) G[W1 + k * H + i] += x[k] * d1[i];
    }
    *loss = L / (double)n;
    return 0;
}
