"""Host-slice execution of large batches (Plan.Forward / Inverse with host
slices, gpu/plan.go:100-143): chunks of whole transforms pipelined over two
streams.  Results must equal the device-resident execution of the same plan
bit for bit (each transform runs the same kernel), from pageable and from
page-locked host memory, with a tail chunk shorter than the others; rows on
both sides of every chunk boundary and the rows of the tail chunk are also
checked against the oracle."""
import ctypes

import numpy as np
import pytest
import torch

import algofft_hip as A
from algofft_hip import synth
import pyoracle as O
from helpers import c64_tol, rel_l2

# rows of the 4100-row batch around the boundaries of its 512-row chunks,
# and the four rows of the tail chunk
CHUNK_ROWS = sorted({0, 1, 4095, 4096, 4097, 4099} | {k * 512 + d for k in range(1, 8) for d in (-1, 0)})


pytestmark = pytest.mark.gpu


def _oracle_rows(x, rows, inverse=False):
    return O.fft_batch(x[rows].astype(np.complex128), inverse=inverse)


def _device_result(plan, direction, x):
    xt = torch.from_numpy(x).to("cuda:0")
    yt = torch.empty(plan.io_len(direction)[1], dtype=xt.dtype, device="cuda:0")
    s = torch.cuda.Stream()
    plan.exec_device(direction, yt.data_ptr(), xt.data_ptr(), s.cuda_stream)
    s.synchronize()
    return yt.cpu().numpy()


class Pinned:
    """Page-locked host array from afft_host_alloc."""

    def __init__(self, ctx, shape, dtype):
        self.arr_bytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        p = ctypes.c_void_p()
        A.check(A.lib().afft_host_alloc(ctx.h, self.arr_bytes, ctypes.byref(p)))
        self.p = p
        buf = (ctypes.c_char * self.arr_bytes).from_address(p.value)
        self.a = np.frombuffer(buf, dtype=dtype).reshape(shape)

    def free(self):
        self.a = None
        A.check(A.lib().afft_host_free(self.p))


@pytest.mark.parametrize("pinned", [False, True])
def test_c2c_batch_pipeline_matches_device(ctx, pinned):
    n, batch = 4096, 4100            # 134 MB each way: 8 chunks of 512 and a tail of 4
    x = synth.uniform_complex(21, (batch, n), np.complex64)
    with A.Plan.from_desc(ctx, (n,), batch=batch) as p:
        want_f = _device_result(p, A.FORWARD, x).reshape(batch, n)
        if pinned:
            src, dst = Pinned(ctx, (batch, n), np.complex64), Pinned(ctx, (batch, n), np.complex64)
            src.a[:] = x
            p.forward(dst.a, src.a)
            got = dst.a.copy()
            p.inverse(dst.a, src.a)
            got_inv = dst.a.copy()
            src.free()
            dst.free()
        else:
            got = np.empty_like(x)
            p.forward(got, x)
            got_inv = np.empty_like(x)
            p.inverse(got_inv, x)
        want_i = _device_result(p, A.INVERSE, x).reshape(batch, n)
    assert np.array_equal(got, want_f)
    assert np.array_equal(got_inv, want_i)
    assert rel_l2(got[CHUNK_ROWS], _oracle_rows(x, CHUNK_ROWS)) <= c64_tol(n)
    assert rel_l2(got_inv[CHUNK_ROWS], _oracle_rows(x, CHUNK_ROWS, inverse=True)) <= c64_tol(n)


def test_r2c_batch_pipeline_matches_device(ctx):
    n, batch = 4096, 8200            # 134 MB of reals in
    x = synth.uniform_real(4, (batch, n), np.float32)
    with A.Plan.from_desc(ctx, (n,), kind=A.R2C, batch=batch) as p:
        xt = torch.from_numpy(x).to("cuda:0")
        yt = torch.empty((batch, n // 2 + 1), dtype=torch.complex64, device="cuda:0")
        s = torch.cuda.Stream()
        p.exec_device(A.FORWARD, yt.data_ptr(), xt.data_ptr(), s.cuda_stream)
        s.synchronize()
        got = np.empty((batch, n // 2 + 1), np.complex64)
        p.forward(got, x)
    assert np.array_equal(got, yt.cpu().numpy())
    # chunk boundaries and the tail chunk against the oracle's real-FFT
    # restatement (plan_real_generic.go:142-228); chunks hold 16 MiB of the
    # larger side: 1023 rows of 2049 bins, then a tail of 16 rows
    chunk = (16 << 20) // max(4 * n, 8 * (n // 2 + 1))
    rows = sorted({0, batch - 1} | {k * chunk + d for k in range(1, batch // chunk + 1) for d in (-1, 0)})
    ref = np.stack([O.rfft(x[r].astype(np.float64)) for r in rows])
    assert rel_l2(got[rows], ref) <= c64_tol(n)


def test_small_batches_take_the_direct_path(ctx):
    # below the pipeline threshold the single staged copy is used; same bits
    n, batch = 1024, 64
    x = synth.uniform_complex(8, (batch, n), np.complex64)
    with A.Plan.from_desc(ctx, (n,), batch=batch) as p:
        got = np.empty_like(x)
        p.forward(got, x)
        want = _device_result(p, A.FORWARD, x).reshape(batch, n)
    assert np.array_equal(got, want)
