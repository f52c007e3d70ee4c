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
