"""Convolution and correlation on the device against the reference's own
known answer (convolve_test.go:14) and the oracle's Convolve restatement
(convolve.go:9-59); correlation as convolution with b reversed and
conjugated (correlate.go:13-33)."""
import ctypes

import numpy as np
import pytest
import torch

import algofft_hip as A
from algofft_hip import synth
import pyoracle as O
from helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype,tol", [(np.complex64, 1e-4), (np.complex128, 1e-12)])
def test_convolve_known_answer(ctx, dtype, tol):
    got = A.convolve(ctx, np.array([1, 2, 3], dtype), np.array([4, 5], dtype))
    assert np.max(np.abs(got - np.array([4, 13, 22, 15]))) < tol


@pytest.mark.parametrize("na,nb", [(7, 5), (1000, 33), (4096, 4096), (1009, 17), (1, 1)])
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_convolve_and_correlate_match_oracle(ctx, na, nb, dtype):
    a = synth.uniform_complex(na, (na,), dtype)
    b = synth.uniform_complex(nb + 1, (nb,), dtype)
    # complex128: the oracle restates Convolve with F = na+nb-1, which for
    # prime F runs Bluestein on the reference's unreduced chirp (~1e-12 off
    # at F = 8191), so it is held to the reference's own 1e-11
    # (convolve_complex128_test.go:52); numpy's direct convolution to 1e-13
    tol = 2e-6 * np.log2(na + nb + 2) if dtype == np.complex64 else 1e-11
    exact = 2e-6 * np.log2(na + nb + 2) if dtype == np.complex64 else 1e-13
    got = A.convolve(ctx, a, b)
    ref = O.convolve(a.astype(np.complex128), b.astype(np.complex128))
    assert rel_l2(got, ref) <= tol
    assert rel_l2(got, np.convolve(a.astype(np.complex128), b.astype(np.complex128))) <= exact
    corr = A.convolve(ctx, a, b, correlate=True)
    ref_c = O.convolve(a.astype(np.complex128), np.conj(b[::-1]).astype(np.complex128))
    assert rel_l2(corr, ref_c) <= tol


def test_convolve_errors(ctx):
    lib = A.lib()
    a = np.zeros(4, np.complex64)
    out = np.zeros(3, np.complex64)
    p = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    assert lib.afft_convolve_host(ctx.h, 0, 0, p(out), 3, p(a), 4, p(a), 4) == A.ErrLengthMismatch.status
    assert lib.afft_convolve_host(ctx.h, 0, 0, p(out), 3, p(a), 0, p(a), 4) == A.ErrInvalidLength.status
    assert lib.afft_convolve_host(ctx.h, 0, 0, None, 3, p(a), 4, p(a), 4) == A.ErrNilSlice.status


def test_batched_device_convolution(ctx):
    batch, na, nb = 64, 300, 45
    a = synth.uniform_complex(1, (batch, na), np.complex64)
    b = synth.uniform_complex(2, (batch, nb), np.complex64)
    h = ctypes.c_void_p()
    A.check(A.lib().afft_conv_create(ctx.h, A.COMPLEX64, na, nb, batch, 0, ctypes.byref(h)))
    ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    td = torch.zeros((batch, na + nb - 1), dtype=torch.complex64, device="cuda")
    s = torch.cuda.Stream()
    A.check(A.lib().afft_conv_exec(h, ctypes.c_void_p(td.data_ptr()), ctypes.c_void_p(ta.data_ptr()),
                                   ctypes.c_void_p(tb.data_ptr()), ctypes.c_void_p(s.cuda_stream)))
    s.synchronize()
    got = td.cpu().numpy()
    for i in (0, 31, 63):
        assert rel_l2(got[i], O.convolve(a[i].astype(np.complex128), b[i].astype(np.complex128))) <= 2e-5
    A.check(A.lib().afft_conv_destroy(h))


# ---- ConvolveReal (convolve_real.go:9-80) ------------------------------------

@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_convolve_real_known_answer(ctx, dtype):
    """convolve_real_test.go:10-29: [1 2 3] * [4 5] = [4 13 22 15] (to 1e-3
    there; to rounding here)."""
    got = A.convolve(ctx, np.array([1, 2, 3], dtype), np.array([4, 5], dtype))
    assert got.dtype == dtype
    assert np.max(np.abs(got - np.array([4, 13, 22, 15]))) < (1e-5 if dtype == np.float32 else 1e-12)


def test_convolve_real_gaussian_kernel(ctx):
    """convolve_real_test.go:31-56: a box of ones at 8..23 in 32 samples
    smoothed by the normalised Gaussian of radius 2, sigma 1."""
    signal = np.zeros(32, np.float32)
    signal[8:24] = 1
    x = np.arange(-2, 3, dtype=np.float64)
    k = np.exp(-x * x / 2.0)
    kernel = (k.astype(np.float32) * np.float32(1.0 / k.sum())).astype(np.float32)
    want = np.convolve(signal.astype(np.float64), kernel.astype(np.float64))
    got = A.convolve(ctx, signal, kernel)
    assert np.max(np.abs(got - want)) < 1e-6
    assert abs(float(got.sum()) - 16.0) < 1e-4        # the kernel sums to one


@pytest.mark.parametrize("na,nb", [(9, 6), (1000, 33), (4096, 4096), (1009, 17), (1, 1), (2, 1)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_convolve_real_matches_direct(ctx, na, nb, dtype):
    """Random real inputs against numpy's direct float64 convolution and the
    oracle's ConvolveReal (convolve_real_test.go:58-86 checks the naive sum
    to 1e-3)."""
    a = synth.uniform_real(na, (na,), dtype)
    b = synth.uniform_real(nb + 1, (nb,), dtype)
    got = A.convolve(ctx, a, b)
    ref = np.convolve(a.astype(np.float64), b.astype(np.float64))
    tol = 2e-6 * np.log2(na + nb + 2) if dtype == np.float32 else 1e-13
    assert rel_l2(got, ref) <= tol
    # the oracle's ConvolveReal restatement in FP64 (power-of-two padding,
    # where the device pads to an even 7-smooth length: same convolution)
    assert rel_l2(got, O.convolve_real(a.astype(np.float64), b.astype(np.float64))) <= tol


def test_convolve_real_errors(ctx):
    """convolve_real_test.go:88-120: nil, empty and length-mismatch inputs."""
    lib = A.lib()
    a = np.zeros(4, np.float32)
    out = np.zeros(7, np.float32)
    p = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    R = A.CONV_REAL
    assert lib.afft_convolve_host(ctx.h, 0, R, None, 7, p(a), 4, p(a), 4) == A.ErrNilSlice.status
    assert lib.afft_convolve_host(ctx.h, 0, R, p(out), 7, None, 4, p(a), 4) == A.ErrNilSlice.status
    assert lib.afft_convolve_host(ctx.h, 0, R, p(out), 7, p(a), 4, None, 4) == A.ErrNilSlice.status
    assert lib.afft_convolve_host(ctx.h, 0, R, p(out), 0, p(a), 0, p(a), 4) == A.ErrInvalidLength.status
    assert lib.afft_convolve_host(ctx.h, 0, R, p(out), 0, p(a), 4, p(a), 0) == A.ErrInvalidLength.status
    assert lib.afft_convolve_host(ctx.h, 0, R, p(out), 1, p(a), 2, p(a), 2) == A.ErrLengthMismatch.status
    h = ctypes.c_void_p()
    assert lib.afft_conv_create(ctx.h, 0, 4, 4, 1, 7, ctypes.byref(h)) == A.ErrInvalidArgument.status


def test_batched_device_real_convolution(ctx):
    batch, na, nb = 32, 777, 65
    a = synth.uniform_real(3, (batch, na), np.float32)
    b = synth.uniform_real(4, (batch, nb), np.float32)
    h = ctypes.c_void_p()
    A.check(A.lib().afft_conv_create(ctx.h, A.COMPLEX64, na, nb, batch, A.CONV_REAL, ctypes.byref(h)))
    try:
        at, bt = torch.from_numpy(a).to("cuda:0"), torch.from_numpy(b).to("cuda:0")
        out = torch.empty((batch, na + nb - 1), dtype=torch.float32, device="cuda:0")
        s = torch.cuda.Stream()
        A.check(A.lib().afft_conv_exec(h, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(at.data_ptr()),
                                       ctypes.c_void_p(bt.data_ptr()), ctypes.c_void_p(s.cuda_stream)))
        s.synchronize()
        got = out.cpu().numpy()
    finally:
        A.check(A.lib().afft_conv_destroy(h))
    for i in range(batch):
        ref = np.convolve(a[i].astype(np.float64), b[i].astype(np.float64))
        assert rel_l2(got[i], ref) <= 2e-6 * np.log2(na + nb)
