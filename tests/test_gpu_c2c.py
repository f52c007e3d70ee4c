"""Parity of the HIP complex transforms with the oracle (FP64 restatement of
algo-fft's Plan.Forward/Inverse/ForwardBatch, plan.go:332-571).

Budgets against the FP64 oracle: complex64 rel L2 <= 2e-6*log2(N);
complex128 rel L2 <= 1e-13.
"""
import numpy as np
import pytest

import algofft_hip as A
from algofft_hip import synth
import pyoracle as O
from helpers import C128_TOL, c64_tol, rel_l2

pytestmark = pytest.mark.gpu

POW2 = [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384]


def _run(ctx, x, inverse=False, **kw):
    prec = A.COMPLEX64 if x.dtype == np.complex64 else A.COMPLEX128
    batch, n = x.shape
    with A.Plan.from_desc(ctx, (n,), precision=prec, batch=batch, **kw) as p:
        out = np.empty_like(x)
        (p.inverse if inverse else p.forward)(out, x)
        return out, p.describe()


@pytest.mark.parametrize("n", POW2 + [32768, 65536, 1 << 18, 1 << 20])
def test_c64_forward_pow2(ctx, n):
    batch = max(1, min(64, (1 << 20) // n))
    x = synth.uniform_complex(11 + n, (batch, n), np.complex64)
    got, algo = _run(ctx, x)
    ref = O.fft_batch(x.astype(np.complex128))
    err = rel_l2(got, ref)
    print(f"n={n} batch={batch} algo={algo} relL2={err:.3e}")
    assert err <= c64_tol(n), (n, err, algo)


@pytest.mark.parametrize("n", POW2[:-1] + [16384, 65536, 1 << 20])
def test_c128_forward_pow2(ctx, n):
    batch = max(1, min(16, (1 << 18) // n))
    x = synth.uniform_complex(23 + n, (batch, n), np.complex128)
    got, algo = _run(ctx, x)
    ref = O.fft_batch(x)
    err = rel_l2(got, ref)
    print(f"n={n} batch={batch} algo={algo} relL2={err:.3e}")
    assert err <= C128_TOL, (n, err, algo)


@pytest.mark.parametrize("n", [8, 256, 1024, 4096, 16384, 65536])
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_inverse_and_roundtrip(ctx, n, dtype):
    batch = 8
    x = synth.uniform_complex(5, (batch, n), dtype)
    X, _ = _run(ctx, x)
    y, _ = _run(ctx, X, inverse=True)
    ref_inv = O.fft_batch(X.astype(np.complex128), inverse=True)
    tol = c64_tol(n) if dtype == np.complex64 else C128_TOL
    assert rel_l2(y, ref_inv) <= tol
    assert rel_l2(y, x) <= 2 * tol


def test_batch_1024x256_equals_loop(ctx):
    # N=1024, batch 256, contiguous: the batch result equals the per-transform
    # results (plan_batch_test.go:13-146)
    n, batch = 1024, 256
    x = synth.uniform_complex(1, (batch, n), np.complex64)
    got, _ = _run(ctx, x)
    with A.Plan.new(ctx, n, A.COMPLEX64) as p:
        for i in (0, 17, 255):
            one = np.empty(n, np.complex64)
            p.forward(one, x[i])
            assert np.array_equal(one, got[i])
    ref = O.fft_batch(x.astype(np.complex128))
    assert rel_l2(got, ref) <= c64_tol(n)


def test_c1_inverse_1024x256_and_round_trip(ctx):
    # the plain CPU-runnable workload in the other direction: InverseBatch of
    # 256 transforms of 1024 points against the oracle's inverse (1/N
    # scaling, plan.go:545-571), and the forward-inverse round trip
    n, batch = 1024, 256
    x = synth.uniform_complex(1, (batch, n), np.complex64)
    got, _ = _run(ctx, x, inverse=True)
    ref = O.fft_batch(x.astype(np.complex128), inverse=True)
    err = rel_l2(got, ref)
    print(f"inverse n={n} batch={batch} relL2={err:.3e}")
    assert err <= c64_tol(n)
    X, _ = _run(ctx, x)
    back, _ = _run(ctx, X, inverse=True)
    assert rel_l2(back, x) <= 2 * c64_tol(n)


def test_in_place(ctx):
    n = 4096
    x = synth.uniform_complex(3, (4, n), np.complex64)
    with A.Plan.from_desc(ctx, (n,), batch=4, flags=A.FLAG_INPLACE) as p:
        buf = x.copy()
        p.forward(buf, buf)
    got, _ = _run(ctx, x)
    assert np.array_equal(buf, got)


def test_known_signals(ctx):
    # plan_signals_test.go:8-157: impulse -> ones, constant -> N at DC,
    # bin-k exponential -> N at k, alternating -> N at Nyquist
    n = 4096
    x = np.zeros((4, n), np.complex64)
    x[0, 0] = 1
    x[1, :] = 1
    k = 37
    x[2] = np.exp(2j * np.pi * k * np.arange(n) / n).astype(np.complex64)
    x[3] = (-1.0) ** np.arange(n)
    got, _ = _run(ctx, x)
    e = np.zeros((4, n), np.complex128)
    e[0, :] = 1
    e[1, 0] = n
    e[2, k] = n
    e[3, n // 2] = n
    assert np.max(np.abs(got - e)) <= 1e-6 * n * np.log2(n)


def test_strided_columns(ctx):
    # ForwardStrided over the columns of a row-major matrix (plan_strided.go:35-99)
    n, cols = 1024, 48
    a = synth.uniform_complex(9, (n, cols), np.complex64)
    with A.Plan.from_desc(ctx, (n,), batch=cols, istride=cols, idist=1, ostride=cols, odist=1) as p:
        out = np.empty_like(a)
        p.forward(out, a)
    ref = O.fft_batch(np.ascontiguousarray(a.T).astype(np.complex128)).T
    assert rel_l2(out, ref) <= c64_tol(n)


def test_single_strided(ctx):
    n, stride = 256, 300
    buf = synth.uniform_complex(4, ((n - 1) * stride + 1,), np.complex128)
    with A.Plan.from_desc(ctx, (n,), precision=A.COMPLEX128, istride=stride, ostride=stride) as p:
        out = np.zeros_like(buf)
        p.forward(out, buf)
    ref = O.fft_strided(buf, n, stride)
    idx = np.arange(n) * stride
    assert rel_l2(out[idx], ref[idx]) <= C128_TOL


@pytest.mark.parametrize("dims,dtype", [((4096, 64), np.complex64), ((16384, 16), np.complex64),
                                        ((2048, 32), np.complex128), ((3, 4096, 20), np.complex64),
                                        ((8192, 17), np.complex128)])
def test_long_column_axes(ctx, dims, dtype):
    # axes too long for 8 adjacent columns per block run as a column
    # four-step (two coalesced passes); checked against the PlanND oracle
    prec = A.COMPLEX64 if dtype == np.complex64 else A.COMPLEX128
    x = synth.uniform_complex(7, dims, dtype)
    with A.Plan.from_desc(ctx, dims, precision=prec) as p:
        algo = p.describe()
        out = np.empty_like(x)
        p.forward(out, x)
        back = np.empty_like(x)
        p.inverse(back, out)
    n = int(np.prod(dims))
    tol = c64_tol(n) if dtype == np.complex64 else C128_TOL
    assert "colfourstep" in algo, algo
    assert rel_l2(out, O.fftn(x.astype(np.complex128))) <= tol
    assert rel_l2(back, x) <= 2 * tol


@pytest.mark.parametrize("dims", [(64, 128), (512, 512), (16, 32, 64), (8, 4, 16, 32)])
def test_nd_c2c(ctx, dims):
    x = synth.uniform_complex(2, dims, np.complex64)
    with A.Plan.from_desc(ctx, dims) as p:
        out = np.empty_like(x)
        p.forward(out, x)
        back = np.empty_like(x)
        p.inverse(back, out)
    ref = O.fftn(x.astype(np.complex128))
    n = int(np.prod(dims))
    assert rel_l2(out, ref) <= c64_tol(n)
    assert rel_l2(back, x) <= 2 * c64_tol(n)


TUNED_SPLITS = {(20, A.COMPLEX64): 512, (23, A.COMPLEX64): 4096, (24, A.COMPLEX64): 8192,
                (18, A.COMPLEX128): 1024, (19, A.COMPLEX128): 256, (23, A.COMPLEX128): 4096}


@pytest.mark.parametrize("lg,prec", sorted(TUNED_SPLITS))
def test_tuned_fourstep_splits(ctx, lg, prec):
    """The lengths whose default four-step split comes from the tuning
    records: the plan takes that split, matches numpy's FP64 FFT and its
    inverse returns the input."""
    n = 1 << lg
    split = TUNED_SPLITS[(lg, prec)]
    cdt = np.complex64 if prec == A.COMPLEX64 else np.complex128
    x = synth.uniform_complex(lg, (n,), cdt)
    with A.Plan.from_desc(ctx, (n,), precision=prec) as p:
        assert p.describe() == f"hip_fourstep{split}x{n // split}"
        X = np.empty_like(x)
        p.forward(X, x)
        y = np.empty_like(x)
        p.inverse(y, X)
    tol = c64_tol(n) if prec == A.COMPLEX64 else C128_TOL
    assert rel_l2(X, np.fft.fft(x.astype(np.complex128))) <= tol
    assert rel_l2(y, x) <= 2 * tol
