"""Pins the oracle (oracle/afft_oracle.c, the C restatement of algo-fft) before
it is trusted as the parity checker.

Checks, each chosen so that a plausible mistake fails at least one:
  * the reference's own known-answer tests: twiddles for n=4 and n=8
    (internal/math/twiddle_test.go:70-120), trivial DFTs
    (internal/reference/dft_test.go:118-158), the convolution [4,13,22,15]
    (convolve_test.go:14), constant / bin / Nyquist signals
    (plan_signals_test.go:8-157);
  * brute-force long-double DFT for every route at small N;
  * numpy.fft golden vectors committed under tests/golden;
  * invariants: Parseval, linearity, shift theorem, round trip, batch == loop,
    real-input Hermitian symmetry, separability of the N-D transform;
  * planted mistakes (or_set_fault) are caught by these checks.
"""
import os

import numpy as np
import pytest

import pyoracle as O
from algofft_hip import synth
from helpers import rel_l2

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "numpy_fft_vectors.npz"))


def test_twiddle_known_answers():
    e4 = np.array([1, -1j, -1, 1j])
    s = np.sqrt(2) / 2
    e8 = np.array([1, s - s * 1j, -1j, -s - s * 1j, -1, -s + s * 1j, 1j, s + s * 1j])
    assert np.max(np.abs(O.twiddles(4, np.complex64) - e4)) < 1e-6
    assert np.max(np.abs(O.twiddles(8, np.complex64) - e8)) < 1e-6
    assert np.max(np.abs(O.twiddles(4) - e4)) < 1e-14
    assert np.max(np.abs(O.twiddles(8) - e8)) < 1e-14


@pytest.mark.parametrize("fn", [O.fft, O.dft_naive])
def test_trivial_dfts(fn):
    assert np.allclose(fn(np.array([1, 0, 0, 0], np.complex128)), [1, 1, 1, 1], atol=1e-12)
    assert np.allclose(fn(np.array([1, 1, 1, 1], np.complex128)), [4, 0, 0, 0], atol=1e-12)
    assert np.allclose(fn(np.array([1, 1], np.complex128)), [2, 0], atol=1e-12)


@pytest.mark.parametrize("dtype,tol", [(np.complex64, 1e-4), (np.complex128, 1e-12)])
def test_convolution_known_answer(dtype, tol):
    got = O.convolve(np.array([1, 2, 3], dtype), np.array([4, 5], dtype))
    assert np.max(np.abs(got - np.array([4, 13, 22, 15]))) < tol


@pytest.mark.parametrize("dtype,tol", [(np.float32, 1e-4), (np.float64, 1e-12)])
def test_real_convolution_known_answers(dtype, tol):
    """ConvolveReal (convolve_real.go:9-80) against the reference's own known
    answer (convolve_real_test.go:10-29), its Gaussian smoothing case
    (:31-56) and direct summation (:58-86, here numpy's np.convolve in FP64)
    at lengths whose padding is and is not already a power of two."""
    got = O.convolve_real(np.array([1, 2, 3], dtype), np.array([4, 5], dtype))
    assert np.max(np.abs(got - np.array([4, 13, 22, 15]))) < tol
    signal = np.zeros(32, dtype)
    signal[8:24] = 1
    x = np.arange(-2, 3, dtype=np.float64)
    k = np.exp(-x * x / 2.0)
    kernel = (k / k.sum()).astype(dtype)
    ref = np.convolve(signal.astype(np.float64), kernel.astype(np.float64))
    assert np.max(np.abs(O.convolve_real(signal, kernel) - ref)) < (1e-5 if dtype == np.float32 else 1e-13)
    for na, nb in ((9, 6), (1, 1), (512, 513), (1000, 33)):
        a = synth.uniform_real(na, (na,), dtype)
        b = synth.uniform_real(nb + 1, (nb,), dtype)
        ref = np.convolve(a.astype(np.float64), b.astype(np.float64))
        err = np.linalg.norm(O.convolve_real(a, b) - ref) / np.linalg.norm(ref)
        assert err < (2e-6 * np.log2(na + nb + 2) if dtype == np.float32 else 1e-13), (na, nb, err)


ROUTE_SIZES = {1: [16, 256, 1024], 2: [4096, 1 << 18], 3: [6, 60, 1000, 3125], 4: [7, 49, 1009, 4093]}


@pytest.mark.parametrize("route,n", [(r, n) for r, ns in ROUTE_SIZES.items() for n in ns])
def test_routes_match_reference_planner(route, n):
    # planner.go:43-78: pow2 -> Stockham or six-step, 2/3/5-smooth -> mixed
    # radix, everything else (7-smooth included) -> Bluestein
    assert O.route(n) == route


@pytest.mark.parametrize("n", [16, 4096, 1000, 1009, 1 << 18])
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_known_signals(n, dtype):
    t = np.arange(n)
    k = 5
    cases = [
        (np.ones(n), np.eye(1, n, 0)[0] * n),
        (np.exp(2j * np.pi * k * t / n), np.eye(1, n, k)[0] * n),
    ]
    if n % 2 == 0:
        cases.append(((-1.0) ** t, np.eye(1, n, n // 2)[0] * n))
    tol = 2e-6 * np.log2(n) if dtype == np.complex64 else 1e-12
    for x, want in cases:
        got = O.fft(x.astype(dtype))
        assert np.max(np.abs(got - want)) / n < tol


def test_naive_dft_against_closed_forms():
    n = 12
    x = np.exp(2j * np.pi * 3 * np.arange(n) / n)
    got = O.dft_naive(x)
    assert np.max(np.abs(got - np.eye(1, n, 3)[0] * n)) < 1e-13
    assert np.allclose(O.dft_naive(got, inverse=True), x, atol=1e-14)


@pytest.mark.parametrize("n", list(range(1, 65)) + [100, 243, 343, 1000, 1009])
def test_brute_force(n):
    x = synth.uniform_complex(n, (n,), np.complex128)
    ref = O.dft_naive(x)
    assert rel_l2(O.fft(x), ref) < 1e-14 * max(1, np.log2(n))
    assert rel_l2(O.fft(ref, inverse=True), x) < 1e-14 * max(1, np.log2(n))
    assert rel_l2(O.fft(x.astype(np.complex64)), O.dft_naive(x.astype(np.complex64))) < 2e-6 * max(1, np.log2(n))


def test_golden_c2c():
    for key in GOLD.files:
        if key.startswith("c2c_x_"):
            n = key[6:]
            x = GOLD[key]
            assert rel_l2(O.fft(x), GOLD["c2c_X_" + n]) < 1e-14 * max(1, np.log2(len(x))), n


def test_golden_r2c_and_nd():
    for key in GOLD.files:
        if key.startswith("r2c_x_"):
            n = key[6:]
            x = GOLD[key]
            assert rel_l2(O.rfft(x[None])[0], GOLD["r2c_X_" + n]) < 1e-14, n
        if key.startswith("nd_x_"):
            x = GOLD[key]
            assert rel_l2(O.fftn(x), GOLD["nd_X_" + key[5:]]) < 1e-14, key


def test_generator_streams_agree():
    # the C stream (or_rng_u01) and the Python stream are the same numbers
    assert np.array_equal(O.rng_u01(0, 16), GOLD["rng_seed0_first16"])
    assert np.array_equal(O.rng_u01(1234, 16), GOLD["rng_seed1234_first16"])
    assert np.array_equal(synth.u01(1234, 1000), O.rng_u01(1234, 1000))


@pytest.mark.parametrize("n", [256, 1000, 1009, 4096])
def test_invariants(n):
    x = synth.uniform_complex(1, (n,), np.complex128)
    y = synth.uniform_complex(2, (n,), np.complex128)
    X, Y = O.fft(x), O.fft(y)
    # Parseval (precision_test.go:93-191)
    assert abs(np.sum(np.abs(X) ** 2) / n - np.sum(np.abs(x) ** 2)) / np.sum(np.abs(x) ** 2) < 1e-13
    # linearity (properties_test.go:15)
    assert rel_l2(O.fft(2.5 * x - 1j * y), 2.5 * X - 1j * Y) < 1e-13
    # shift theorem (properties_test.go:303)
    s = 7
    Xs = O.fft(np.roll(x, s))
    assert rel_l2(Xs, X * np.exp(-2j * np.pi * s * np.arange(n) / n)) < 1e-12
    # round trip (precision_test.go:207-250)
    assert rel_l2(O.fft(X, inverse=True), x) < 1e-14
    # batch == loop (plan_batch_test.go:13-146)
    b = np.stack([x, y])
    B = O.fft_batch(b, nthreads=2)
    assert np.array_equal(B[0], X) and np.array_equal(B[1], Y)


@pytest.mark.parametrize("n", [8, 64, 1000, 2018])
def test_real_matches_complex(n):
    x = synth.uniform_real(3, (n,), np.float64)
    X = O.rfft(x[None])[0]
    full = O.fft(x.astype(np.complex128))
    # Hermitian symmetry and R2C == first half of C2C (properties_test.go:469)
    assert rel_l2(X, full[: n // 2 + 1]) < 1e-14
    assert rel_l2(full[1:][::-1], np.conj(full[1:])) < 1e-13
    assert rel_l2(O.irfft(X[None], n)[0], x) < 1e-14


def test_nd_separable():
    x = synth.uniform_complex(4, (6, 10), np.complex128)
    rows = O.fft_batch(x)
    cols = O.fft_batch(np.ascontiguousarray(rows.T)).T
    assert rel_l2(O.fftn(x), cols) < 1e-14
    xr = synth.uniform_real(4, (4, 6, 8), np.float64)
    assert rel_l2(O.rfftn(xr), np.fft.rfftn(xr)) < 1e-14


def test_strided_equals_contiguous():
    n, stride = 100, 3
    buf = synth.uniform_complex(5, (n * stride,), np.complex128)
    got = O.fft_strided(buf, n, stride)
    assert rel_l2(got[::stride], O.fft(buf[::stride].copy())) < 1e-15


def test_bluestein_reference_chirp_deviation():
    # the reference's unreduced chirp (bluestein.go:23-28) loses accuracy as
    # k^2 grows; the exact reduction does not
    n = 100003
    exact = O.chirp(n, exact=True)
    refch = O.chirp(n, exact=False)
    k = np.arange(n, dtype=np.float64)
    # float128 reference phase from the integer reduction
    kk = (np.arange(n, dtype=np.int64) ** 2) % (2 * n)
    truth = np.exp(-1j * np.pi * kk.astype(np.float64) / n)
    assert np.max(np.abs(exact - truth)) < 4e-15
    assert np.max(np.abs(refch - truth)) > 1e-13
    del k


@pytest.mark.parametrize("fault,check", [
    (1, lambda: rel_l2(O.fft(GOLD["c2c_x_1024"]), GOLD["c2c_X_1024"]) < 1e-12),
    (2, lambda: rel_l2(O.fft(GOLD["c2c_x_1000"]), GOLD["c2c_X_1000"]) < 1e-12),
    (3, lambda: rel_l2(O.fft(GOLD["c2c_x_1009"]), GOLD["c2c_X_1009"]) < 1e-12),
    (4, lambda: rel_l2(O.fftn(GOLD["nd_x_3x5x8"]), GOLD["nd_X_3x5x8"]) < 1e-12),
])
def test_planted_mistakes_are_caught(fault, check):
    assert check()
    O.set_fault(fault)
    try:
        assert not check()
    finally:
        O.set_fault(0)
    assert check()
