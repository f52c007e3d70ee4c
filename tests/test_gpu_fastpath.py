"""Zero-validation fast paths: FastPlan (plan_fast.go:21-133) and
Plan.Clone / Executor (plan.go:1300, executor.go:6-43).  A fast plan runs the
same kernel as the checked plan, so its output is bit-identical; in place is
allowed; clones transform independently on their own streams."""
import threading

import numpy as np
import pytest
import torch

import algofft_hip as A
from algofft_hip import synth
import pyoracle as O
from helpers import C128_TOL, c64_tol, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
@pytest.mark.parametrize("n", [1, 2, 16, 256, 1024, 4096, 8192])
def test_fast_plan_matches_plan_bits(ctx, n, dtype):
    prec = A.COMPLEX64 if dtype == np.complex64 else A.COMPLEX128
    batch = 3
    x = synth.uniform_complex(n + 1, (batch, n), dtype)
    xt = torch.from_numpy(x).to("cuda:0")
    s = torch.cuda.Stream()
    with A.FastPlan(ctx, n, batch, prec, s.cuda_stream) as fp, A.Plan.from_desc(ctx, (n,), precision=prec,
                                                                                  batch=batch) as p:
        assert len(fp) == n
        yf, yp = torch.empty_like(xt), torch.empty_like(xt)
        with torch.cuda.stream(s):
            fp.forward(yf.data_ptr(), xt.data_ptr())
        p.exec_device(A.FORWARD, yp.data_ptr(), xt.data_ptr(), s.cuda_stream)
        s.synchronize()
        assert torch.equal(yf, yp)
        tol = c64_tol(max(n, 2)) if dtype == np.complex64 else C128_TOL
        assert rel_l2(yf.cpu().numpy(), O.fft_batch(x.astype(np.complex128))) <= tol
        # in place, inverse: back to the input
        z = yf.clone()
        with torch.cuda.stream(s):
            fp.inverse(z.data_ptr(), z.data_ptr())
        s.synchronize()
        assert rel_l2(z.cpu().numpy(), x) <= 2 * tol


def test_fast_plan_errors(ctx):
    with pytest.raises(A.ErrInvalidLength):
        A.FastPlan(ctx, 1000)
    with pytest.raises(A.ErrNotImplemented):
        A.FastPlan(ctx, 1 << 16)       # four-step: more than one launch
    with pytest.raises(A.ErrNotImplemented):
        A.FastPlan(ctx, 16384, precision=A.COMPLEX128)


def test_clones_run_concurrently(ctx):
    n, batch = 4096, 64
    with A.Plan.from_desc(ctx, (n,), batch=batch) as p:
        clones = [p.clone() for _ in range(4)]
        xs = [synth.uniform_complex(40 + i, (batch, n), np.complex64) for i in range(4)]
        outs = [np.empty_like(x) for x in xs]

        def work(i):
            for _ in range(3):
                clones[i].forward(outs[i], xs[i])

        th = [threading.Thread(target=work, args=(i,)) for i in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        for i in range(4):
            want = np.empty_like(xs[i])
            p.forward(want, xs[i])
            assert np.array_equal(outs[i], want)
            assert clones[i].describe() == p.describe()
            clones[i].close()
