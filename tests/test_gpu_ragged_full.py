"""The mixed-size workload at its full size: 4096 complex64 transforms,
lengths log-uniform in [16, 65536] (60 % powers of two, 30 % 7-smooth, 10 %
primes through Bluestein), 34.1 M points in 734 distinct lengths, built by
synth.ragged_sizes(1234, 4096) exactly as the benchmark builds it.  Every job
of the graph-replayed persistent launch is checked against the oracle's
Plan.Forward / Plan.Inverse restatement (plan.go:332-419), forward and
inverse, as the reference's batch tests check every transform of a batch
(plan_batch_test.go:13-146).  The plain per-size comparator (one batched plan
per distinct length, on a copy sorted by length) is checked the same way."""
import numpy as np
import pytest
import torch

import algofft_hip as A
from algofft_hip import synth
import pyoracle as O
from helpers import c64_tol, rel_l2

pytestmark = pytest.mark.gpu

SIZES = synth.ragged_sizes(1234, 4096)


@pytest.fixture(scope="module")
def workload():
    total = sum(SIZES)
    offs = np.concatenate([[0], np.cumsum(SIZES)[:-1]]).astype(np.int64)
    x = synth.uniform_complex(4096, (total,), np.complex64)
    xd = x.astype(np.complex128)
    ref = {}
    for inverse in (False, True):
        r = np.empty(total, np.complex128)
        for n, o in zip(SIZES, offs):
            r[o:o + n] = O.fft(xd[o:o + n], inverse=inverse)
        ref[inverse] = r
    return x, offs, ref


def _worst(sizes, offs, got, ref):
    worst, where = 0.0, None
    for n, o in zip(sizes, offs):
        ratio = rel_l2(got[o:o + n], ref[o:o + n]) / c64_tol(n)
        if ratio > worst:
            worst, where = ratio, n
    return worst, where


def test_full_ragged_set_graph_forward_and_inverse(ctx, workload):
    x, offs, ref = workload
    assert len(SIZES) == 4096 and len(set(SIZES)) == 734 and sum(SIZES) == 34061541
    with A.Ragged(ctx, SIZES, A.COMPLEX64) as rg:
        assert list(rg.offsets()) == list(offs)
        xt = torch.from_numpy(x).to("cuda:0")
        s = torch.cuda.Stream()
        for direction, inverse in ((A.FORWARD, False), (A.INVERSE, True)):
            yt = torch.zeros_like(xt)
            rg.exec_graph(direction, yt.data_ptr(), xt.data_ptr(), s.cuda_stream)
            s.synchronize()
            worst, n = _worst(SIZES, offs, yt.cpu().numpy(), ref[inverse])
            print(f"{'inverse' if inverse else 'forward'}: worst rel_l2/budget {worst:.4f} (n={n})", flush=True)
            assert worst <= 1.0, (worst, n)


def test_full_ragged_set_per_size_comparator(ctx, workload):
    x, offs, ref = workload
    order = sorted(range(len(SIZES)), key=lambda i: (SIZES[i], i))
    ssizes = [SIZES[i] for i in order]
    soffs = np.concatenate([[0], np.cumsum(ssizes)[:-1]]).astype(np.int64)
    xs = np.concatenate([x[offs[i]:offs[i] + SIZES[i]] for i in order])
    rs = np.concatenate([ref[False][offs[i]:offs[i] + SIZES[i]] for i in order])
    with A.Ragged(ctx, ssizes, A.COMPLEX64) as rg:
        xt = torch.from_numpy(xs).to("cuda:0")
        s = torch.cuda.Stream()
        yp = torch.zeros_like(xt)
        rg.exec_per_size(A.FORWARD, yp.data_ptr(), xt.data_ptr(), s.cuda_stream)
        yg = torch.zeros_like(xt)
        rg.exec_graph(A.FORWARD, yg.data_ptr(), xt.data_ptr(), s.cuda_stream)
        s.synchronize()
        for name, y in (("per-size", yp), ("graph, sorted set", yg)):
            worst, n = _worst(ssizes, soffs, y.cpu().numpy(), rs)
            print(f"{name}: worst rel_l2/budget {worst:.4f} (n={n})", flush=True)
            assert worst <= 1.0, (name, worst, n)
