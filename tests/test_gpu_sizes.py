"""Non-power-of-two sizes: the runtime mixed-radix path (2/3/5/7-smooth) and
the Bluestein chain (all other sizes), against the oracle's mixed-radix and
Bluestein restatements (internal/fft/mixedradix.go:45-358,
plan_bluestein.go:7-58) on the same seeded inputs.

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

SMOOTH = [3, 5, 6, 7, 9, 12, 15, 30, 49, 100, 243, 360, 1000, 2187, 2401, 3000, 3125, 6000, 8000, 12000, 59049]
PRIMEISH = [11, 13, 17, 31, 101, 251, 509, 1009, 4093, 8191, 8209, 65521, 10007 * 3]


def _plan_run(ctx, x, inverse=False, flags=0, **kw):
    prec = A.COMPLEX64 if x.dtype == np.complex64 else A.COMPLEX128
    batch, n = x.shape
    with A.Plan.from_desc(ctx, (n,), precision=prec, batch=batch, flags=flags, **kw) as p:
        out = np.empty_like(x)
        (p.inverse if inverse else p.forward)(out, x)
        return out, p.describe()


@pytest.mark.parametrize("n", SMOOTH + PRIMEISH)
def test_c64_forward(ctx, n):
    batch = max(1, min(16, 200000 // n))
    x = synth.uniform_complex(n, (batch, n), np.complex64)
    got, algo = _plan_run(ctx, x)
    ref = O.fft_batch(x.astype(np.complex128))
    err = rel_l2(got, ref)
    print(f"n={n} algo={algo} relL2={err:.2e}")
    assert err <= c64_tol(n), (n, err, algo)


@pytest.mark.parametrize("n", [3, 7, 100, 2187, 3125, 12000, 13, 509, 8191, 65521])
def test_c128_forward_inverse(ctx, n):
    batch = max(1, min(8, 100000 // n))
    x = synth.uniform_complex(7 * n, (batch, n), np.complex128)
    got, algo = _plan_run(ctx, x)
    ref = O.fft_batch(x)
    assert rel_l2(got, ref) <= C128_TOL, (n, rel_l2(got, ref), algo)
    back, _ = _plan_run(ctx, got, inverse=True)
    ref_b = O.fft_batch(got, inverse=True)
    assert rel_l2(back, ref_b) <= C128_TOL
    assert rel_l2(back, x) <= 2 * C128_TOL


@pytest.mark.parametrize("n", [6, 360, 3000, 12000, 17, 1009, 8209])
def test_c64_inverse(ctx, n):
    x = synth.uniform_complex(3 * n, (4, n), np.complex64)
    got, algo = _plan_run(ctx, x, inverse=True)
    ref = O.fft_batch(x.astype(np.complex128), inverse=True)
    assert rel_l2(got, ref) <= c64_tol(n), (n, rel_l2(got, ref), algo)


def test_length_one(ctx):
    x = synth.uniform_complex(1, (5, 1), np.complex64)
    got, algo = _plan_run(ctx, x)
    assert np.array_equal(got, x)
    back, _ = _plan_run(ctx, x, inverse=True)
    assert np.array_equal(back, x)


@pytest.mark.parametrize("n", [1009, 8209])
def test_reference_chirp_flag(ctx, n):
    # the unreduced chirp of bluestein.go:23-28 is selectable; at these sizes
    # both chirps agree to rounding, and both match the oracle's matching form
    x = synth.uniform_complex(5, (2, n), np.complex128)
    got, _ = _plan_run(ctx, x, flags=A.FLAG_REF_CHIRP)
    ref = O.fft_batch(x, exact_chirp=False)
    assert rel_l2(got, ref) <= 1e-12


@pytest.mark.parametrize("n,stride", [(360, 7), (17, 40), (8209, 3), (12000, 2)])
def test_strided_mixed_and_bluestein(ctx, n, stride):
    buf = synth.uniform_complex(n + stride, ((n - 1) * stride + 1,), np.complex64)
    with A.Plan.from_desc(ctx, (n,), precision=A.COMPLEX64, istride=stride, ostride=stride) as p:
        out = np.zeros_like(buf)
        p.forward(out, buf)
    ref = O.fft_strided(buf.astype(np.complex128), n, stride)
    idx = np.arange(n) * stride
    assert rel_l2(out[idx], ref[idx]) <= c64_tol(n)


@pytest.mark.parametrize("dims", [(6, 10), (7, 9, 11), (3, 5, 64), (100, 1), (1, 17)])
def test_nd_mixed(ctx, dims):
    x = synth.uniform_complex(11, dims, np.complex128)
    with A.Plan.from_desc(ctx, dims, precision=A.COMPLEX128) as p:
        out = np.empty_like(x)
        p.forward(out, x)
        back = np.empty_like(x)
        p.inverse(back, out)
    assert rel_l2(out, O.fftn(x)) <= C128_TOL
    assert rel_l2(back, x) <= 2 * C128_TOL


def test_interleaved_columns_mixed(ctx):
    n, cols = 360, 33
    a = synth.uniform_complex(9, (n, cols), np.complex64)
    with A.Plan.from_desc(ctx, (n,), batch=cols, istride=cols, idist=1, ostride=cols, odist=1) as p:
        out = np.empty_like(a)
        p.forward(out, a)
    ref = O.fft_batch(np.ascontiguousarray(a.T).astype(np.complex128)).T
    assert rel_l2(out, ref) <= c64_tol(n)


# Contiguous smooth rows held whole in one LDS buffer (one transform per
# block, passes in place): complex64 from about 3000 to 16384 points,
# complex128 from about 2600 to 8192 (the two-buffer kernel takes shorter
# rows, and rows where it keeps more threads resident).
IP_C64 = [3500, 4000, 4500, 5184, 6144, 7000, 9000, 10240, 13122, 15000, 16000, 16128]
IP_C128 = [2592, 3000, 4000, 5000, 7168, 8000]


@pytest.mark.parametrize("prec,n", [(A.COMPLEX64, n) for n in IP_C64] + [(A.COMPLEX128, n) for n in IP_C128])
def test_in_place_rows_forward_inverse(ctx, prec, n):
    dt = np.complex64 if prec == A.COMPLEX64 else np.complex128
    tol = c64_tol(n) if prec == A.COMPLEX64 else C128_TOL
    batch = 5
    x = synth.uniform_complex(13 * n, (batch, n), dt)
    got, algo = _plan_run(ctx, x)
    assert algo == f"hip_mixed{n}r", algo
    ref = O.fft_batch(x.astype(np.complex128))
    assert rel_l2(got, ref) <= tol, (n, rel_l2(got, ref))
    # the inverse carries 1/N; FLAG_NORMALIZE puts 1/N on the forward instead
    back, _ = _plan_run(ctx, got, inverse=True)
    ref_b = O.fft_batch(got.astype(np.complex128), inverse=True)
    assert rel_l2(back, ref_b) <= tol, (n, rel_l2(back, ref_b))
    assert rel_l2(back, x) <= 2 * tol
    scaled, _ = _plan_run(ctx, x, flags=A.FLAG_NORMALIZE)
    assert rel_l2(scaled, ref / n) <= tol


@pytest.mark.parametrize("dims", [(3, 12000), (12000, 4), (2, 3, 7000)])
def test_in_place_rows_in_nd_plans(ctx, dims):
    # the contiguous axis runs as in-place rows; a long strided axis is
    # gathered into rows first
    x = synth.uniform_complex(17, dims, np.complex64)
    with A.Plan.from_desc(ctx, dims, precision=A.COMPLEX64) as p:
        out = np.empty_like(x)
        p.forward(out, x)
        algo = p.describe()
    assert rel_l2(out, O.fftn(x.astype(np.complex128))) <= c64_tol(int(np.prod(dims))), algo


# Smooth four-step plans, whose runtime-radix passes run in place in one LDS
# buffer holding several columns (or rows) per block: column counts that are
# not a multiple of the columns per block, one point per thread (n1 = 3), a
# compile-time column pass beside a runtime row pass (n1 = 8), and a row
# pass above 16 points per thread (complex64 n2 = 6000, on the two-buffer
# kernel; complex128 factors stop at 4096 points).
FOURSTEP_SPLITS = [(3000, 3), (3000, 250), (1000, 8), (2187, 27), (5000, 2), (36000, 0), (100000, 0), (250000, 0)]
FOURSTEP_CASES = ([(A.COMPLEX64, n, s) for n, s in FOURSTEP_SPLITS + [(12000, 2)]]
                  + [(A.COMPLEX128, n, s) for n, s in FOURSTEP_SPLITS])


@pytest.mark.parametrize("prec,n,split", FOURSTEP_CASES)
def test_mixed_fourstep_passes(ctx, prec, n, split):
    dt = np.complex64 if prec == A.COMPLEX64 else np.complex128
    batch = max(1, min(5, 300000 // n))
    x = synth.uniform_complex(11 * n + split, (batch, n), dt)
    with A.Plan.from_desc(ctx, (n,), precision=prec, batch=batch, strategy=A.STRATEGY_SIXSTEP, split=split) as p:
        algo = p.describe()
        got = np.empty_like(x)
        p.forward(got, x)
        back = np.empty_like(x)
        p.inverse(back, got)
    assert "mixed4step" in algo, algo
    if split:
        assert f"mixed4step{split}x{n // split}" in algo, algo
    tol = c64_tol(n) if prec == A.COMPLEX64 else C128_TOL
    err = rel_l2(got, O.fft_batch(x.astype(np.complex128)))
    print(f"n={n} split={split} algo={algo} relL2={err:.2e}")
    assert err <= tol, (n, split, err, algo)
    assert rel_l2(back, O.fft_batch(got.astype(np.complex128), inverse=True)) <= tol, (n, split, algo)


# Strided smooth axes of N-D plans above 128 points run on the in-place
# column kernel (below, on the two-buffer one): 16 to 256 threads a column,
# transform counts that are not a multiple of the columns per block, an
# axis with no unit-stride neighbour (the middle axis of a 3-D shape), both
# directions, against the oracle's PlanND restatement.
ND_SMOOTH = [(1200, 15), (300, 7, 5), (5, 1000, 3), (2187, 4), (3125, 3), (6, 250, 250)]


@pytest.mark.parametrize("dims", ND_SMOOTH)
@pytest.mark.parametrize("prec", [A.COMPLEX64, A.COMPLEX128])
def test_strided_smooth_axes(ctx, dims, prec):
    dt = np.complex64 if prec == A.COMPLEX64 else np.complex128
    x = synth.uniform_complex(23 + len(dims), dims, dt)
    with A.Plan.from_desc(ctx, dims, precision=prec) as p:
        algo = p.describe()
        got = np.empty_like(x)
        p.forward(got, x)
        back = np.empty_like(x)
        p.inverse(back, got)
    n = int(np.prod(dims))
    tol = c64_tol(n) if prec == A.COMPLEX64 else C128_TOL
    err = rel_l2(got, O.fftn(x.astype(np.complex128)))
    print(f"dims={dims} algo={algo} relL2={err:.2e}")
    assert err <= tol, (dims, err, algo)
    assert rel_l2(back, O.fftn(got.astype(np.complex128), inverse=True)) <= tol, (dims, algo)


@pytest.mark.parametrize("prec,n", [(A.COMPLEX64, 18000), (A.COMPLEX64, 24000), (A.COMPLEX64, 100000),
                                    (A.COMPLEX128, 9000), (A.COMPLEX128, 36000)])
def test_default_smooth_fourstep_split(ctx, prec, n):
    # the built-in routing's split equals its Python mirror (the planner's
    # heuristic baseline) and the plan matches the oracle
    from algofft_hip import planner
    dt = np.complex64 if prec == A.COMPLEX64 else np.complex128
    x = synth.uniform_complex(31 + n, (2, n), dt)
    got, algo = _plan_run(ctx, x)
    s = planner.fourstep_split(n, prec)
    assert algo == f"hip_mixed4step{s}x{n // s}", (algo, s)
    tol = c64_tol(n) if prec == A.COMPLEX64 else C128_TOL
    assert rel_l2(got, O.fft_batch(x.astype(np.complex128))) <= tol, algo
