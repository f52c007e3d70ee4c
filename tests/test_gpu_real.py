"""Real transforms: PlanRealT Forward/Inverse (plan_real_generic.go:142-355),
PlanReal2D/3D (plan_real_3d.go:145-226) and their ForwardFull/InverseFull
(plan_real_2d.go:231-389, plan_real_3d.go:241-419) against the oracle
restatements (R2C by packing + recombination, C2R by the closed-form repack,
the full spectrum by the complex N-D transform).

Budgets against the FP64 oracle: float32 rel L2 <= 2e-6*log2(N);
float64 rel L2 <= 1e-13.
"""
import numpy as np
import pytest

import algofft_hip as A
from algofft_hip import synth
import pyoracle as O
from helpers import C128_TOL, c64_tol, rel_l2

pytestmark = pytest.mark.gpu

SIZES = [2, 4, 6, 16, 32, 64, 512, 1000, 1024, 2018, 4096, 8192, 32768, 65536]


def _rplan(ctx, dims, prec, batch=1, **kw):
    return A.Plan.from_desc(ctx, dims, precision=prec, kind=A.R2C, batch=batch, **kw)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("prec", [A.COMPLEX64, A.COMPLEX128])
def test_r2c_c2r_1d(ctx, n, prec):
    rdt = np.float32 if prec == A.COMPLEX64 else np.float64
    cdt = np.complex64 if prec == A.COMPLEX64 else np.complex128
    batch = max(1, min(8, 100000 // n))
    x = synth.uniform_real(n, (batch, n), rdt)
    with _rplan(ctx, (n,), prec, batch) as p:
        X = np.empty((batch, n // 2 + 1), cdt)
        p.forward(X, x)
        y = np.empty_like(x)
        p.inverse(y, X)
        algo = p.describe()
    ref = O.rfft(x.astype(np.float64))
    tol = c64_tol(n) if prec == A.COMPLEX64 else C128_TOL
    err = rel_l2(X, ref)
    print(f"n={n} algo={algo} r2c relL2={err:.2e}")
    assert err <= tol, (n, err, algo)
    ref_inv = O.irfft(X.astype(np.complex128), n)
    assert rel_l2(y, ref_inv) <= tol
    assert rel_l2(y, x) <= 2 * tol


def test_c2r_against_oracle_on_hermitian_input(ctx):
    n = 2048
    x = synth.uniform_real(3, (4, n), np.float64)
    X = np.fft.rfft(x, axis=1)
    with _rplan(ctx, (n,), A.COMPLEX128, 4) as p:
        y = np.empty((4, n))
        p.inverse(y, X)
    assert rel_l2(y, O.irfft(X, n)) <= C128_TOL


def test_invalid_spectrum(ctx):
    # plan_real_generic.go:298-316: DC / Nyquist with imaginary part > 1e-4
    n = 64
    X = np.zeros((1, n // 2 + 1), np.complex64)
    X[0, 0] = 1 + 1e-3j
    with _rplan(ctx, (n,), A.COMPLEX64) as p:
        y = np.empty((1, n), np.float32)
        with pytest.raises(A.ErrInvalidSpectrum):
            p.inverse(y, X)
        X[0, 0] = 1 + 1e-5j   # within tolerance: accepted, imaginary part ignored
        p.inverse(y, X)
    ref = O.irfft(np.array([[1] + [0] * (n // 2)], np.complex128), n)
    assert rel_l2(y, ref) <= c64_tol(n)


def test_odd_length_rejected(ctx):
    # newPlanRealTWithFeatures: n < 2 or odd -> ErrInvalidLength (plan_real_generic.go:54-56)
    for n in (1, 7):
        with pytest.raises(A.ErrInvalidLength):
            _rplan(ctx, (n,), A.COMPLEX64)


@pytest.mark.parametrize("dims", [(16, 32), (6, 10), (8, 16, 32), (5, 7, 12), (64, 64, 64)])
@pytest.mark.parametrize("prec", [A.COMPLEX64, A.COMPLEX128])
def test_real_nd(ctx, dims, prec):
    rdt = np.float32 if prec == A.COMPLEX64 else np.float64
    cdt = np.complex64 if prec == A.COMPLEX64 else np.complex128
    x = synth.uniform_real(17, dims, rdt)
    with _rplan(ctx, dims, prec) as p:
        X = np.empty(dims[:-1] + (dims[-1] // 2 + 1,), cdt)
        p.forward(X, x)
        y = np.empty_like(x)
        p.inverse(y, X)
    ref = O.rfftn(x.astype(np.float64))
    n = int(np.prod(dims))
    tol = c64_tol(n) if prec == A.COMPLEX64 else C128_TOL
    assert rel_l2(X, ref) <= tol
    assert rel_l2(X, np.fft.rfftn(x.astype(np.float64))) <= tol
    assert rel_l2(y, x) <= 2 * tol


def test_real_batch_with_distances(ctx):
    n, batch = 256, 5
    rd, cd = 300, 140   # padded rows
    x = synth.uniform_real(2, (batch * rd,), np.float32)
    with _rplan(ctx, (n,), A.COMPLEX64, batch, idist=rd, odist=cd) as p:
        X = np.zeros(batch * cd, np.complex64)
        p.forward(X, x)
    for b in range(batch):
        ref = O.rfft(x[b * rd:b * rd + n].astype(np.float64)[None])[0]
        assert rel_l2(X[b * cd:b * cd + n // 2 + 1], ref) <= c64_tol(n)


def test_forward_normalized_and_unitary(ctx):
    n = 512
    x = synth.uniform_real(8, (2, n), np.float64)
    ref = O.rfft(x)
    for flag, s in ((A.FLAG_NORMALIZE, 1.0 / n), (A.FLAG_UNITARY, 1.0 / np.sqrt(n))):
        with _rplan(ctx, (n,), A.COMPLEX128, 2, flags=flag) as p:
            X = np.empty((2, n // 2 + 1), np.complex128)
            p.forward(X, x)
        assert rel_l2(X, ref * s) <= C128_TOL


@pytest.mark.parametrize("dims", [(4, 4, 4), (3, 4, 6), (5, 8), (6, 10), (64, 128), (16,)])
@pytest.mark.parametrize("prec", [A.COMPLEX64, A.COMPLEX128])
def test_forward_full_inverse_full(ctx, dims, prec):
    """PlanReal2D/3D.ForwardFull and InverseFull (plan_real_2d.go:231-276,
    :370-389; plan_real_3d.go:241-285, :392-419), checked as the reference's
    TestPlanReal3D_ForwardFull / _InverseFull (plan_real_3d_test.go:133-228)
    check them -- the first W/2+1 bins of every row are the compact
    spectrum, InverseFull inverts ForwardFull -- and every bin, the redundant
    ones included, against the full complex transform of the same input."""
    rdt = np.float32 if prec == A.COMPLEX64 else np.float64
    cdt = np.complex64 if prec == A.COMPLEX64 else np.complex128
    x = synth.uniform_real(23, dims, rdt)
    hw = dims[-1] // 2 + 1
    with _rplan(ctx, dims, prec) as p:
        X = np.empty(dims[:-1] + (hw,), cdt)
        p.forward(X, x)
        F = np.empty(dims, cdt)
        p.forward_full(F, x)
        y = np.empty_like(x)
        p.inverse_full(y, F)
    n = int(np.prod(dims))
    tol = c64_tol(n) if prec == A.COMPLEX64 else C128_TOL
    assert np.array_equal(F[..., :hw], X)
    assert rel_l2(F, O.fftn(x.astype(np.complex128))) <= tol
    assert rel_l2(F, np.fft.fftn(x.astype(np.float64))) <= tol
    assert rel_l2(y, x) <= 2 * tol


def test_full_spectrum_device_matches_host(ctx):
    """The device-pointer form gives the host-slice form's bits."""
    dims = (32, 48, 64)
    n = int(np.prod(dims))
    x = synth.uniform_real(5, dims, np.float32)
    with _rplan(ctx, dims, A.COMPLEX64) as p:
        F = np.empty(dims, np.complex64)
        p.forward_full(F, x)
        s = A.Stream(ctx)
        bx, bF, by = A.Buffer(ctx, n // 2), A.Buffer(ctx, n), A.Buffer(ctx, n // 2)
        try:
            bx.upload(x.reshape(-1).view(np.complex64))   # n reals as n/2 complex words
            p.exec_full_device(A.FORWARD, bF.ptr, bx.ptr, s.native)
            p.exec_full_device(A.INVERSE, by.ptr, bF.ptr, s.native)
            s.synchronize()
            Fd = np.empty(n, np.complex64)
            bF.download(Fd)
            yd = np.empty(n // 2, np.complex64)
            by.download(yd)
        finally:
            for b in (bx, bF, by):
                b.close()
            s.close()
    assert np.array_equal(Fd.reshape(dims), F)
    assert rel_l2(yd.view(np.float32).reshape(dims), x) <= 2 * c64_tol(n)


def test_full_spectrum_errors(ctx):
    x = np.zeros(32, np.float32)
    with _rplan(ctx, (4, 8), A.COMPLEX64) as p:
        with pytest.raises(A.ErrLengthMismatch):
            p.forward_full(np.empty(31, np.complex64), x)
        with pytest.raises(A.ErrLengthMismatch):
            p.inverse_full(np.empty(32, np.float32), np.zeros(33, np.complex64))
    with _rplan(ctx, (4, 8), A.COMPLEX64, batch=2) as p:
        with pytest.raises(A.ErrNotImplemented):
            p.forward_full(np.empty(64, np.complex64), np.zeros(64, np.float32))
    with A.Plan.from_desc(ctx, (4, 8)) as p:
        with pytest.raises(A.ErrNotImplemented):
            p.forward_full(np.empty(32, np.complex64), x)
