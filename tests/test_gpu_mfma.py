"""MFMA form (AFFT_FLAG_MFMA): radix-16 passes as FP16 hi/lo matrix products
with FP32 accumulation, checked against the FP64 oracle at the complex64
budget rel L2 <= 2e-6*log2(N), on uniform data and on the precision workload
(8 tones, amplitudes 1e-3..1e1, noise)."""
import numpy as np
import pytest

import algofft_hip as A
from algofft_hip import synth
import pyoracle as O
from helpers import c64_tol, rel_l2

pytestmark = pytest.mark.gpu


def _run(ctx, x, flags, inverse=False):
    batch, n = x.shape
    with A.Plan.from_desc(ctx, (n,), batch=batch, flags=flags) as p:
        y = np.empty_like(x)
        (p.inverse if inverse else p.forward)(y, x)
        return y, p.describe()


@pytest.mark.parametrize("n", [256, 512, 1024, 2048, 4096, 8192, 16384, 65536, 1 << 17, 1 << 18, 1 << 20,
                               1 << 21, 1 << 22, 1 << 24])
@pytest.mark.parametrize("inverse", [False, True])
def test_mfma_matches_oracle(ctx, n, inverse):
    batch = max(1, min(64, (1 << 21) // n))
    x = synth.uniform_complex(n, (batch, n), np.complex64)
    y, algo = _run(ctx, x, A.FLAG_MFMA, inverse)
    assert "mfma" in algo, algo
    ref = O.fft_batch(x.astype(np.complex128), inverse=inverse)
    err = rel_l2(y, ref)
    plain, _ = _run(ctx, x, 0, inverse)
    print(f"n={n} inverse={inverse} mfma relL2={err:.3e} plain relL2={rel_l2(plain, ref):.3e}")
    assert err <= c64_tol(n)


@pytest.mark.parametrize("n", [256, 1024, 2048, 4096, 16384, 65536, 1 << 20])
def test_mfma_on_high_dynamic_range_tones(ctx, n):
    rows = 32 if n <= 4096 else 4
    x, bins, amps = synth.sinusoids(77, n, rows)
    y, _ = _run(ctx, x, A.FLAG_MFMA)
    ref = O.fft_batch(x.astype(np.complex128))
    assert rel_l2(y, ref) <= c64_tol(n)
    # the strongest tone of every row lands in its bin
    for b in range(rows):
        j = int(np.argmax(amps[b]))
        assert abs(y[b, bins[b, j]]) >= 0.5 * n * amps[b, j]


def test_mfma_strided_four_step_slices(ctx):
    # a batch of 65536-point rows at a row distance larger than n
    n, batch, dist = 65536, 3, 65536 + 512
    x = synth.uniform_complex(9, (batch, dist), np.complex64)
    with A.Plan.from_desc(ctx, (n,), batch=batch, idist=dist, odist=dist, flags=A.FLAG_MFMA) as p:
        y = np.zeros_like(x)
        p.forward(y, x)
    ref = O.fft_batch(x[:, :n].astype(np.complex128))
    assert rel_l2(y[:, :n], ref) <= c64_tol(n)


def test_mfma_per_transform_scale(ctx):
    """Each transform is scaled by its own power of two: rows of very
    different magnitude in one block (1e-30 .. 1e30, and an all-zero row)
    each keep the complex64 budget relative to themselves."""
    n, batch = 256, 16
    x = synth.uniform_complex(41, (batch, n), np.complex64)
    mags = np.array([1e-30, 1e30, 0.0, 1.0, 3e-5, 7e4, 1e-3, 1e3] * 2, np.float64)
    x = (x * mags[:, None]).astype(np.complex64)
    y, algo = _run(ctx, x, A.FLAG_MFMA)
    assert "mfma" in algo
    ref = O.fft_batch(x.astype(np.complex128))
    for b in range(batch):
        if mags[b] == 0:
            assert not np.any(y[b])
        else:
            assert rel_l2(y[b], ref[b]) <= c64_tol(n), (b, mags[b])


def test_mfma_unsupported_shapes_are_reported(ctx):
    for n in (128, 1000, 1 << 25):
        with pytest.raises(A.ErrNotImplemented):
            A.Plan.from_desc(ctx, (n,), flags=A.FLAG_MFMA)
    with pytest.raises(A.ErrNotImplemented):
        A.Plan.from_desc(ctx, (4096,), precision=A.COMPLEX128, flags=A.FLAG_MFMA)


def test_mfma_plans_ignore_wisdom_for_the_plain_kernels():
    """Wisdom ranks the plain kernels: an imported pick of a four-step split
    the MFMA form does not build (2^20 as 512 x 2048), or of the mixed-radix
    kernel for 4096, must not reach an MFMA plan.  The plan keeps the MFMA
    route and its result stays within budget; a plain plan follows the
    wisdom."""
    with A.Context(0) as c2:
        # batch 2 is bucket 1 and batch 16 bucket 2 of the wisdom shape class
        text = f"{1 << 20}:0:{0x950 | (1 << 16)}:hip_sixstep_512:1\n4096:0:{0x950 | (2 << 16)}:hip_mixed:1\n"
        A.check(A.lib().afft_wisdom_import(c2.h, text.encode()))
        for n, plain_algo in ((1 << 20, "fourstep512x2048"), (4096, "mixed4096r")):
            batch = 2 if n > 4096 else 16
            x = synth.uniform_complex(n + 1, (batch, n), np.complex64)
            ref = O.fft_batch(x.astype(np.complex128))
            with A.Plan.from_desc(c2, (n,), batch=batch, flags=A.FLAG_MFMA) as p:
                algo = p.describe()
                y = np.empty_like(x)
                p.forward(y, x)
            assert "mfma" in algo, algo
            assert rel_l2(y, ref) <= c64_tol(n)
            with A.Plan.from_desc(c2, (n,), batch=batch) as p:
                assert plain_algo in p.describe(), p.describe()
        # strategies without an MFMA form are refused outright
        for strat in (A.STRATEGY_MIXED, A.STRATEGY_BLUESTEIN):
            with pytest.raises(A.ErrNotImplemented):
                A.Plan.from_desc(c2, (4096,), flags=A.FLAG_MFMA, strategy=strat)
