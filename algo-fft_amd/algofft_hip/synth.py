"""Seeded synthetic inputs shared by tests and the benchmark.

uniform(): counter-based splitmix64 stream, value i of stream `seed` is
    z = splitmix64(seed * 0xD1342543DE82EF95 + offset + i);  u = (z >> 11) * 2^-53
mapped to [-1, 1) as 2u - 1.  Real parts take even counters, imaginary parts
odd ones.  The C oracle implements the same stream (or_rng_u01).
"""
import numpy as np

_M1 = np.uint64(0x9E3779B97F4A7C15)
_M2 = np.uint64(0xBF58476D1CE4E5B9)
_M3 = np.uint64(0x94D049BB133111EB)
_SEEDMUL = np.uint64(0xD1342543DE82EF95)


def _splitmix64(x):
    with np.errstate(over="ignore"):
        z = x + _M1
        z = (z ^ (z >> np.uint64(30))) * _M2
        z = (z ^ (z >> np.uint64(27))) * _M3
        return z ^ (z >> np.uint64(31))


def u01(seed, count, offset=0):
    """count doubles in [0, 1) from the stream."""
    with np.errstate(over="ignore"):
        base = np.uint64(seed) * _SEEDMUL + np.uint64(offset)
        idx = base + np.arange(count, dtype=np.uint64)
    return (_splitmix64(idx) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniform_complex(seed, shape, dtype=np.complex64):
    """i.i.d. U[-1, 1) real and imaginary parts."""
    n = int(np.prod(shape))
    u = u01(seed, 2 * n) * 2.0 - 1.0
    out = (u[0::2] + 1j * u[1::2]).astype(dtype)
    return out.reshape(shape)


def uniform_real(seed, shape, dtype=np.float32):
    n = int(np.prod(shape))
    return (u01(seed, n) * 2.0 - 1.0).astype(dtype).reshape(shape)


def normal_complex(seed, shape, dtype=np.complex128):
    """N(0, 1) real and imaginary parts by Box-Muller on the stream."""
    n = int(np.prod(shape))
    u = u01(seed, 4 * n)
    u1 = 1.0 - u[0::2]  # (0, 1]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    g = r * np.cos(2 * np.pi * u2)
    return (g[0::2] + 1j * g[1::2]).astype(dtype).reshape(shape)


def sinusoids(seed, n, batch, dtype=np.complex64, tones=8, noise=1e-3):
    """Precision workload: each row is a sum of `tones` complex exponentials at
    seeded integer bins, amplitudes log-uniform in [1e-3, 1e1], uniform phase,
    plus complex N(0, noise) Gaussian noise (dynamic range ~1e4).
    Returns (signal, bins, amps) with bins/amps of shape (batch, tones)."""
    u = u01(seed, 3 * tones * batch).reshape(batch, tones, 3)
    bins = np.minimum((u[..., 0] * n).astype(np.int64), n - 1)
    amps = 10.0 ** (-3.0 + 4.0 * u[..., 1])
    phase = 2 * np.pi * u[..., 2]
    t = np.arange(n, dtype=np.int64)
    sig = np.zeros((batch, n), dtype=np.complex128)
    for j in range(tones):
        # exact phase: (bin * t) mod n before scaling by 2π/n
        ph = ((bins[:, j:j + 1] * t[None, :]) % n).astype(np.float64) * (2 * np.pi / n) + phase[:, j:j + 1]
        sig += amps[:, j:j + 1] * np.exp(1j * ph)
    sig += normal_complex(seed + 1, (batch, n)) * noise
    return sig.astype(dtype), bins, amps


def _is_smooth7(n):
    for q in (2, 3, 5, 7):
        while n % q == 0:
            n //= q
    return n == 1


def _is_prime(n):
    if n < 2:
        return False
    if n % 2 == 0:
        return n == 2
    f = 3
    while f * f <= n:
        if n % f == 0:
            return False
        f += 2
    return True


def ragged_sizes(seed, count=4096, lo=16, hi=65536):
    """Ragged-batch lengths: log-uniform in [lo, hi]; 60% rounded to a power
    of two, 30% to the nearest 2^a 3^b 5^c 7^d, 10% to the nearest prime
    (the Bluestein sizes)."""
    u = u01(seed, 2 * count).reshape(count, 2)
    out = []
    for i in range(count):
        v = int(round(np.exp(np.log(lo) + (np.log(hi) - np.log(lo)) * u[i, 0])))
        v = min(max(v, lo), hi)
        c = u[i, 1]
        if c < 0.6:
            n = 1 << int(round(np.log2(v)))
        elif c < 0.9:
            d = 0
            while True:
                if v - d >= lo and _is_smooth7(v - d):
                    n = v - d
                    break
                if v + d <= hi and _is_smooth7(v + d):
                    n = v + d
                    break
                d += 1
        else:
            d = 0
            while True:
                if _is_prime(v - d) and v - d >= 2:
                    n = v - d
                    break
                if _is_prime(v + d):
                    n = v + d
                    break
                d += 1
        out.append(min(max(n, 2), 1 << 20))
    return out
