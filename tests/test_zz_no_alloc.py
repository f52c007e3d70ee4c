"""afft_plan_exec_device_noalloc: plan execution with the scratch taken from
the library's static device array instead of memory the plan allocates.

Called through the C ABI.  Each case runs a plan with workspace through
afft_plan_exec_device and through the allocation-free entry point on the same
device input: the outputs must be bit-identical (same kernels) and within
the FP64 oracle's budget.  Calls come in pairs, a large workspace first and a
smaller one second, so the second call runs over what the first left in the
shared array."""
import ctypes

import numpy as np
import pytest
import torch

import algofft_hip as A
from algofft_hip import synth
import pyoracle as O
from helpers import C128_TOL, c64_tol, rel_l2

pytestmark = pytest.mark.gpu


def _noalloc(p, direction, dst, src, stream):
    return A.lib().afft_plan_exec_device_noalloc(p.h, direction, ctypes.c_void_p(dst.data_ptr()),
                                                 ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(stream))


def _both(p, direction, x_dev, out_shape, out_dtype):
    """(output of afft_plan_exec_device, output of the no-alloc entry point)."""
    s = torch.cuda.Stream()
    y_ref = torch.empty(out_shape, dtype=out_dtype, device="cuda:0")
    y_na = torch.full(out_shape, float("nan"), dtype=out_dtype, device="cuda:0") if not out_dtype.is_complex \
        else torch.full(out_shape, complex("nan+nanj"), dtype=out_dtype, device="cuda:0")
    torch.cuda.synchronize()
    p.exec_device(direction, y_ref.data_ptr(), x_dev.data_ptr(), s.cuda_stream)
    rc = _noalloc(p, direction, y_na, x_dev, s.cuda_stream)
    assert rc == A.OK, A.lib().afft_last_error()
    s.synchronize()
    return y_ref.cpu().numpy(), y_na.cpu().numpy()


def _c2c_case(ctx, n, batch, prec, seed):
    dt = np.complex64 if prec == A.COMPLEX64 else np.complex128
    x = synth.uniform_complex(seed, (batch, n), dt)
    xd = torch.from_numpy(x).to("cuda:0")
    with A.Plan.from_desc(ctx, (n,), precision=prec, batch=batch) as p:
        ws = p.workspace_bytes()
        assert 0 < ws <= A.lib().afft_noalloc_scratch_bytes(), (n, batch, ws)
        print(f"noalloc c2c n={n} batch={batch} prec={prec} ws={ws} algo={p.describe()}", flush=True)
        y_ref, y_na = _both(p, A.FORWARD, xd, (batch, n), xd.dtype)
        yd = torch.from_numpy(y_na).to("cuda:0")
        z_ref, z_na = _both(p, A.INVERSE, yd, (batch, n), xd.dtype)
    assert np.array_equal(y_na, y_ref)
    assert np.array_equal(z_na, z_ref)
    tol = c64_tol(n) if prec == A.COMPLEX64 else C128_TOL
    assert rel_l2(y_na, O.fft_batch(x.astype(np.complex128))) <= tol
    assert rel_l2(z_na, x) <= 2 * tol


@pytest.mark.parametrize("first,second", [
    ((65536, 48, A.COMPLEX64), (65536, 3, A.COMPLEX64)),       # four-step: TMP0
    ((1 << 18, 16, A.COMPLEX64), (4099, 5, A.COMPLEX128)),     # four-step, then a Bluestein chain
    ((40009, 8, A.COMPLEX128), (20011, 3, A.COMPLEX64)),      # Bluestein chains: TMP0 and TMP1
])
def test_noalloc_matches_exec_device_twice(ctx, first, second):
    _c2c_case(ctx, *first, seed=11)
    _c2c_case(ctx, *second, seed=12)


def test_noalloc_real_nd_inverse_work_buffer(ctx):
    # the inverse of a real 3-D plan runs its outer axes into WORK first
    for dims, prec, seed in (((32, 64, 128), A.COMPLEX128, 5), ((8, 12, 16), A.COMPLEX64, 6)):
        rdt = np.float32 if prec == A.COMPLEX64 else np.float64
        x = synth.uniform_real(seed, dims, rdt)
        X = np.fft.rfftn(x.astype(np.float64)).astype(np.complex64 if prec == A.COMPLEX64 else np.complex128)
        Xd = torch.from_numpy(X).to("cuda:0")
        with A.Plan.from_desc(ctx, dims, precision=prec, kind=A.R2C) as p:
            assert p.workspace_bytes() > 0
            print(f"noalloc c2r dims={dims} prec={prec} ws={p.workspace_bytes()} algo={p.describe()}", flush=True)
            y_ref, y_na = _both(p, A.INVERSE, Xd, dims, torch.from_numpy(x).dtype)
        assert np.array_equal(y_na, y_ref)
        tol = c64_tol(int(np.prod(dims))) if prec == A.COMPLEX64 else C128_TOL
        assert rel_l2(y_na, x) <= 2 * tol


def test_noalloc_rejects_oversize_workspace_and_bad_arguments(ctx):
    cap = A.lib().afft_noalloc_scratch_bytes()
    assert cap == 64 << 20
    n, batch = 1 << 20, 64   # four-step over 512 MiB: workspace beyond the array
    with A.Plan.from_desc(ctx, (n,), batch=batch) as p:
        assert p.workspace_bytes() > cap
        # full-size buffers, although the call is rejected before anything is enqueued
        x = torch.zeros(n * batch, dtype=torch.complex64, device="cuda:0")
        assert _noalloc(p, A.FORWARD, x, x, 0) == A.ErrInvalidLength.status
        assert b"static scratch" in A.lib().afft_last_error()
        assert A.lib().afft_plan_exec_device_noalloc(p.h, 0, ctypes.c_void_p(x.data_ptr()),
                                                     ctypes.c_void_p(x.data_ptr()), None) == A.ErrInvalidArgument.status
        assert A.lib().afft_plan_exec_device_noalloc(p.h, A.FORWARD, None, None, None) == A.ErrNilSlice.status
    assert A.lib().afft_plan_exec_device_noalloc(None, A.FORWARD, None, None, None) == A.ErrNotImplemented.status
