"""Ragged batches (the mixed-size workload): one persistent work-queue launch,
its hipGraph replay, and one plan execution per job, each checked job by
job against the oracle's Plan.Forward/Inverse restatement (plan.go:332-419).
"""
import numpy as np
import pytest
import torch

import algofft_hip as A
from algofft_hip import synth
import pyoracle as O
from helpers import C128_TOL, c64_tol, rel_l2

pytestmark = pytest.mark.gpu


def _check(sizes, offs, x, y, dtype, inverse=False):
    worst = 0.0
    for n, o in zip(sizes, offs):
        ref = O.fft(x[o:o + n].astype(np.complex128), inverse=inverse)
        err = rel_l2(y[o:o + n], ref)
        tol = c64_tol(n) if dtype == np.complex64 else C128_TOL
        assert err <= tol, (n, err)
        worst = max(worst, err / tol)
    return worst


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_ragged_three_forms_agree_with_oracle(ctx, dtype):
    sizes = synth.ragged_sizes(7, 160, 16, 65536) + [1, 2, 3, 5, 7, 17, 4096, 8192, 8191, 65521, 65536, 1 << 18]
    prec = A.COMPLEX64 if dtype == np.complex64 else A.COMPLEX128
    with A.Ragged(ctx, sizes, prec) as rg:
        total = len(rg)
        offs = rg.offsets()
        assert total == sum(sizes)
        x = synth.uniform_complex(3, (total,), dtype)
        dev = torch.device("cuda:0")
        xt = torch.from_numpy(x).to(dev)
        s = torch.cuda.Stream()
        outs = {}
        for form in ("exec", "exec_graph", "exec_per_size", "exec_per_job"):
            yt = torch.zeros_like(xt)
            with torch.cuda.stream(s):
                getattr(rg, form)(A.FORWARD, yt.data_ptr(), xt.data_ptr(), s.cuda_stream)
            s.synchronize()
            outs[form] = yt.cpu().numpy()
            print(form, "worst err/tol", _check(sizes, offs, x, outs[form], dtype))
        # persistent and graph forms run the same kernel: identical bits
        assert np.array_equal(outs["exec"], outs["exec_graph"])
        # inverse round trip through the persistent kernel
        zt = torch.zeros_like(xt)
        yt = torch.from_numpy(outs["exec"]).to(dev)
        with torch.cuda.stream(s):
            rg.exec(A.INVERSE, zt.data_ptr(), yt.data_ptr(), s.cuda_stream)
        s.synchronize()
        _check(sizes, offs, outs["exec"], zt.cpu().numpy(), dtype, inverse=True)


def test_graph_replays_with_eager_work_between(ctx):
    # every replay of the captured (reset kernel + persistent kernel) pair
    # matches the eager launch, with other device work in between
    sizes = synth.ragged_sizes(11, 300)
    with A.Ragged(ctx, sizes, A.COMPLEX64) as rg:
        x = torch.from_numpy(synth.uniform_complex(5, (len(rg),), np.complex64)).to("cuda:0")
        s = torch.cuda.Stream()
        ref = torch.zeros_like(x)
        with torch.cuda.stream(s):
            rg.exec(A.FORWARD, ref.data_ptr(), x.data_ptr(), s.cuda_stream)
        s.synchronize()
        y = torch.zeros_like(x)
        for i in range(4):
            with torch.cuda.stream(s):
                rg.exec_graph(A.FORWARD, y.data_ptr(), x.data_ptr(), s.cuda_stream)
            s.synchronize()
            assert torch.equal(y, ref), i
            junk = torch.randn(1 << 20, device="cuda:0").sum()   # eager work between replays
            torch.cuda.synchronize()
            y.zero_()
            del junk


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_ragged_power_of_two_split_lengths(ctx, dtype):
    """Split jobs at both ends of the power-of-two core lengths: 8192 = 64 x
    128 up to 2^20 = 1024 x 1024, and Bluestein paddings from 8192 (n = 2053)
    to 2^21 = 1024 x 2048 (n = 1048573), beside generic smooth splits."""
    sizes = [1 << 20, 1048573, 8192, 2053, 4099, 1 << 19, 524287, 16384, 100000, 12000]
    prec = A.COMPLEX64 if dtype == np.complex64 else A.COMPLEX128
    with A.Ragged(ctx, sizes, prec) as rg:
        offs = rg.offsets()
        x = synth.uniform_complex(13, (len(rg),), dtype)
        xt = torch.from_numpy(x).to("cuda:0")
        s = torch.cuda.Stream()
        for direction, inverse in ((A.FORWARD, False), (A.INVERSE, True)):
            yt = torch.zeros_like(xt)
            with torch.cuda.stream(s):
                rg.exec_graph(direction, yt.data_ptr(), xt.data_ptr(), s.cuda_stream)
            s.synchronize()
            print("inverse" if inverse else "forward", "worst err/tol",
                  _check(sizes, offs, x, yt.cpu().numpy(), dtype, inverse=inverse))


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_ragged_grouped_small_power_of_two_jobs(ctx, dtype):
    """Whole power-of-two jobs of one length run RAG_PTS/L to a task: full and
    partial groups of every length 16 .. 4096, interleaved with lengths that
    stay on the runtime passes (2, 4, 8, smooth, prime)."""
    rng = np.random.default_rng(21)
    sizes = [16] * 300 + [32] * 130 + [64] * 65 + [128] * 33 + [256] * 17 + [512] * 9 + [1024] * 5 + \
        [2048] * 3 + [4096] * 2 + [2, 4, 8, 3, 1000, 17, 4099]
    sizes = [int(v) for v in rng.permutation(sizes)]
    prec = A.COMPLEX64 if dtype == np.complex64 else A.COMPLEX128
    with A.Ragged(ctx, sizes, prec) as rg:
        offs = rg.offsets()
        x = synth.uniform_complex(17, (len(rg),), dtype)
        xt = torch.from_numpy(x).to("cuda:0")
        s = torch.cuda.Stream()
        for direction, inverse in ((A.FORWARD, False), (A.INVERSE, True)):
            yt = torch.zeros_like(xt)
            with torch.cuda.stream(s):
                rg.exec(direction, yt.data_ptr(), xt.data_ptr(), s.cuda_stream)
            s.synchronize()
            print("inverse" if inverse else "forward", "worst err/tol",
                  _check(sizes, offs, x, yt.cpu().numpy(), dtype, inverse=inverse))


def test_ragged_executions_on_several_streams(ctx):
    """Executions share the queue counters and the split jobs' scratch: one
    enqueued on another stream than the previous one waits for it, so
    back-to-back executions on four streams all give the single-stream bits."""
    sizes = synth.ragged_sizes(23, 200, 16, 70000)
    with A.Ragged(ctx, sizes) as rg:
        xs = [torch.from_numpy(synth.uniform_complex(30 + i, (len(rg),), np.complex64)).to("cuda:0")
              for i in range(4)]
        refs = []
        for x in xs:
            y = torch.zeros_like(x)
            rg.exec(A.FORWARD, y.data_ptr(), x.data_ptr())
            torch.cuda.synchronize()
            refs.append(y)
        streams = [torch.cuda.Stream() for _ in range(4)]
        outs = [torch.zeros_like(x) for x in xs]
        for _ in range(2):
            for s, x, y in zip(streams, xs, outs):
                rg.exec(A.FORWARD, y.data_ptr(), x.data_ptr(), s.cuda_stream)
            for s, x, y in zip(streams, xs, outs):
                rg.exec_graph(A.FORWARD, y.data_ptr(), x.data_ptr(), s.cuda_stream)
        torch.cuda.synchronize()
        for y, r in zip(outs, refs):
            assert torch.equal(y, r)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_ragged_whole_jobs_up_to_the_lds_cap(ctx, dtype):
    """Jobs of up to 16384 points (complex64; 8192 complex128) run whole, in
    one task: power-of-two ones on blocks of L/16 threads, smooth ones on the
    runtime passes with one job per block, and Bluestein jobs grouped by
    padded length (jobs of one group differ in n, chirp and filter), on both
    sides of the cap and beside split jobs."""
    cap = 16384 if dtype == np.complex64 else 8192
    primes_small = [11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97]
    sizes = primes_small * 3 + [2053, 2063, 4099, 4111, 8191, 8209] + \
        [8192, 16384, 4375, 6000, 10000, 15625, 16807, 4097 * 3, 32768, 20000] * 2 + [4096, 2048, 1000]
    rng = np.random.default_rng(5)
    sizes = [int(v) for v in rng.permutation(sizes)]
    prec = A.COMPLEX64 if dtype == np.complex64 else A.COMPLEX128
    with A.Ragged(ctx, sizes, prec) as rg:
        offs = rg.offsets()
        x = synth.uniform_complex(29, (len(rg),), dtype)
        xt = torch.from_numpy(x).to("cuda:0")
        s = torch.cuda.Stream()
        for direction, inverse in ((A.FORWARD, False), (A.INVERSE, True)):
            yt = torch.zeros_like(xt)
            with torch.cuda.stream(s):
                rg.exec_graph(direction, yt.data_ptr(), xt.data_ptr(), s.cuda_stream)
            s.synchronize()
            print("cap", cap, "inverse" if inverse else "forward", "worst err/tol",
                  _check(sizes, offs, x, yt.cpu().numpy(), dtype, inverse=inverse))


@pytest.mark.parametrize("sizes", [[1], [4096], [12000], [1 << 18], [100003], [5003] * 7, [3, 65536, 3]],
                         ids=["one-point", "pow2-whole", "smooth-whole", "pow2-split", "blue-split",
                              "blue-whole-only", "two-chains"])
def test_ragged_sets_with_one_or_two_chains(ctx, sizes):
    # sets whose jobs all fall into one or two of the concurrent chains: the
    # launches of the chains present share the device, the others are absent
    with A.Ragged(ctx, sizes, A.COMPLEX64) as rg:
        offs = rg.offsets()
        x = synth.uniform_complex(29, (len(rg),), np.complex64)
        xt = torch.from_numpy(x).to("cuda:0")
        s = torch.cuda.Stream()
        for form in ("exec", "exec_graph"):
            yt = torch.zeros_like(xt)
            with torch.cuda.stream(s):
                getattr(rg, form)(A.FORWARD, yt.data_ptr(), xt.data_ptr(), s.cuda_stream)
            s.synchronize()
            _check(sizes, offs, x, yt.cpu().numpy(), np.complex64)
