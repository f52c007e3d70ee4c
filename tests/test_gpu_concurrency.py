"""Concurrency, determinism and stress, after the reference's
concurrent_test.go:12-400, plan_fuzz_test.go:9-150 and stress_test.go:17-300:
one plan shared by many host threads, plans created concurrently, mixed
precisions at once, one plan driven from several streams, bit-identical
results across runs, streams, clones and execution paths, seeded random
shapes (the fuzz targets' round trip and determinism), and device memory that
returns to its level after many plans are created and destroyed.
"""
import threading

import numpy as np
import pytest
import torch

import algofft_hip as A
from algofft_hip import synth
import pyoracle as O
from helpers import C128_TOL, c64_tol, rel_l2

pytestmark = pytest.mark.gpu


def _run_threads(fn, count):
    errs = []

    def wrap(i):
        try:
            fn(i)
        except BaseException as e:   # surfaced in the main thread
            errs.append((i, e))

    th = [threading.Thread(target=wrap, args=(i,)) for i in range(count)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    if errs:
        raise AssertionError(f"thread {errs[0][0]}: {errs[0][1]!r}") from errs[0][1]


@pytest.mark.parametrize("threads", [2, 4, 8, 16])
@pytest.mark.parametrize("n", [256, 1024, 4096, 65536])
def test_shared_plan_many_threads(ctx, n, threads):
    """One plan, each thread with its own buffers, 100 forward/inverse cycles
    per thread (concurrent_test.go:12-84).  The first forward of every thread
    matches the oracle; the round trips end near the input."""
    iters = 100 if n <= 4096 else 10
    xs = [synth.uniform_complex(1000 * n + i, (n,), np.complex64) for i in range(threads)]
    first = [None] * threads
    last = [None] * threads
    with A.Plan.new(ctx, n) as p:
        def work(i):
            src = xs[i].copy()
            dst = np.empty_like(src)
            for it in range(iters):
                p.forward(dst, src)
                if it == 0:
                    first[i] = dst.copy()
                p.inverse(src, dst)
            last[i] = src
        _run_threads(work, threads)
    for i in range(threads):
        assert rel_l2(first[i], O.fft(xs[i].astype(np.complex128))) <= c64_tol(n)
        assert rel_l2(last[i], xs[i]) <= iters * c64_tol(n)


def test_concurrent_plan_creation(ctx):
    """Eight threads create and run plans of every kind at once
    (concurrent_test.go:149-211)."""
    shapes = [(256,), (4096,), (1 << 16,), (3000,), (1009,), (64, 96), (8, 16, 32), (12000,), (24000,)]
    inputs = [synth.uniform_complex(7 + i, (2,) + s, np.complex64) for i, s in enumerate(shapes)]
    want = []
    for s, x in zip(shapes, inputs):
        flat = x.reshape(2, -1).astype(np.complex128)
        if len(s) == 1:
            want.append(O.fft_batch(flat))
        else:
            want.append(np.stack([O.fftn(r.reshape(s)).reshape(-1) for r in flat]))
    got = [None] * len(shapes)

    def work(i):
        for _ in range(3):
            with A.Plan.from_desc(ctx, shapes[i], batch=2) as p:
                out = np.empty_like(inputs[i])
                p.forward(out, inputs[i])
        got[i] = out.reshape(2, -1)
    _run_threads(work, len(shapes))
    for i, s in enumerate(shapes):
        n = int(np.prod(s))
        assert rel_l2(got[i], want[i]) <= c64_tol(n), s


def test_concurrent_mixed_precisions(ctx):
    """complex64 and complex128 plans of the same sizes run at once from
    alternating threads (concurrent_test.go:274-345)."""
    sizes = [1024, 3000, 8209, 1 << 16]
    jobs = [(n, dt) for n in sizes for dt in (np.complex64, np.complex128)]
    res = [None] * len(jobs)

    def work(i):
        n, dt = jobs[i]
        x = synth.uniform_complex(n + 3, (2, n), dt)
        prec = A.COMPLEX64 if dt == np.complex64 else A.COMPLEX128
        with A.Plan.from_desc(ctx, (n,), precision=prec, batch=2) as p:
            for _ in range(5):
                y = np.empty_like(x)
                p.forward(y, x)
                z = np.empty_like(x)
                p.inverse(z, y)
        res[i] = (x, y, z)
    _run_threads(work, len(jobs))
    for (n, dt), (x, y, z) in zip(jobs, res):
        tol = c64_tol(n) if dt == np.complex64 else C128_TOL
        assert rel_l2(y, O.fft_batch(x.astype(np.complex128))) <= tol
        assert rel_l2(z, x) <= 2 * tol


@pytest.mark.parametrize("n", [1 << 16, 24000, 8209])
def test_one_plan_from_several_streams(ctx, n):
    """A plan with workspace (four-step, mixed four-step, Bluestein) enqueued
    on four streams in turn: each stream waits for the previous execution's
    use of the workspace, so every result is the single-stream result."""
    batch = 4
    xs = [torch.from_numpy(synth.uniform_complex(n + i, (batch, n), np.complex64)).to("cuda:0") for i in range(4)]
    ys = [torch.empty_like(x) for x in xs]
    streams = [torch.cuda.Stream() for _ in range(4)]
    with A.Plan.from_desc(ctx, (n,), batch=batch) as p:
        assert p.workspace_bytes() > 0
        ref = []
        for x in xs:
            y = torch.empty_like(x)
            p.exec_device(A.FORWARD, y.data_ptr(), x.data_ptr())
            A.check(A.lib().afft_plan_synchronize(p.h))
            ref.append(y)
        torch.cuda.synchronize()
        for _ in range(3):
            for s, x, y in zip(streams, xs, ys):
                p.exec_device(A.FORWARD, y.data_ptr(), x.data_ptr(), s.cuda_stream)
        torch.cuda.synchronize()
    for y, r in zip(ys, ref):
        assert torch.equal(y, r)


@pytest.mark.parametrize("n", [16, 1024, 4096, 1 << 16, 1 << 20, 3000, 12000, 24000, 1009, 100003])
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_bit_identical_across_runs_streams_and_paths(ctx, n, dtype):
    """FuzzDeterministicForward (plan_fuzz_test.go:52-90), made strict: the
    same input gives the same bits on repeated runs, on another stream, on a
    clone and through the host-slice path."""
    prec = A.COMPLEX64 if dtype == np.complex64 else A.COMPLEX128
    batch = 2
    x = synth.uniform_complex(n + 5, (batch, n), dtype)
    xt = torch.from_numpy(x).to("cuda:0")
    s2 = torch.cuda.Stream()
    with A.Plan.from_desc(ctx, (n,), precision=prec, batch=batch) as p:
        outs = []
        for st in (None, None, s2.cuda_stream):
            y = torch.empty_like(xt)
            p.exec_device(A.FORWARD, y.data_ptr(), xt.data_ptr(), st)
            torch.cuda.synchronize()
            A.check(A.lib().afft_plan_synchronize(p.h))
            outs.append(y.cpu().numpy())
        with p.clone() as c:
            yc = np.empty_like(x)
            c.forward(yc, x)
        yh = np.empty_like(x)
        p.forward(yh, x)
    for o in outs[1:] + [yc, yh]:
        assert np.array_equal(o, outs[0])


def test_ragged_bit_identical_across_launch_forms(ctx):
    """The persistent launch, its graph and repeated replays give the same
    bits: every job's arithmetic is independent of which block takes it."""
    sizes = synth.ragged_sizes(9, count=300, lo=16, hi=20000)
    with A.Ragged(ctx, sizes) as r:
        x = torch.from_numpy(synth.uniform_complex(9, (len(r),), np.complex64)).to("cuda:0")
        outs = []
        for form in ("exec", "exec_graph", "exec_graph", "exec"):
            y = torch.empty_like(x)
            getattr(r, form)(A.FORWARD, y.data_ptr(), x.data_ptr())
            torch.cuda.synchronize()
            outs.append(y.cpu())
    for o in outs[1:]:
        assert torch.equal(o, outs[0])


def _random_shapes(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        kind = rng.integers(0, 4)
        if kind == 0:
            n = int(1 << rng.integers(0, 18))                 # powers of two
        elif kind == 1:
            e = rng.integers(0, 6, size=4)                    # 7-smooth, <= 151200
            n = int(2 ** e[0] * 3 ** min(e[1], 3) * 5 ** min(e[2], 2) * 7 ** min(e[3], 1))
        else:
            n = int(rng.integers(1, 70000))                   # anything
        batch = int(rng.integers(1, 5))
        dt = np.complex64 if rng.integers(0, 2) == 0 else np.complex128
        out.append((n, batch, dt))
    return out


@pytest.mark.parametrize("chunk", range(4))
def test_seeded_random_shapes(ctx, chunk):
    """FuzzRoundTripComplex64 / FuzzNoPanicValidInput (plan_fuzz_test.go:9-115)
    over seeded random lengths 1..70000 in both precisions: forward against
    the oracle, the round trip, and in-place execution."""
    for n, batch, dt in _random_shapes(100 + chunk, 20):
        prec = A.COMPLEX64 if dt == np.complex64 else A.COMPLEX128
        tol = c64_tol(max(n, 2)) if dt == np.complex64 else C128_TOL
        x = synth.uniform_complex(n * 7 + batch, (batch, n), dt)
        with A.Plan.from_desc(ctx, (n,), precision=prec, batch=batch) as p:
            y = np.empty_like(x)
            p.forward(y, x)
            z = y.copy()
            p.inverse(z, z)
        assert rel_l2(y, O.fft_batch(x.astype(np.complex128))) <= tol, (n, batch, dt)
        assert rel_l2(z, x) <= 2 * tol, (n, batch, dt)


@pytest.mark.parametrize("chunk", range(2))
def test_seeded_random_real_shapes(ctx, chunk):
    """FuzzRoundTripReal (plan_fuzz_test.go:117-150): even lengths, real
    forward against the oracle and back."""
    rng = np.random.default_rng(500 + chunk)
    for _ in range(15):
        n = 2 * int(rng.integers(1, 35000))
        prec = A.COMPLEX64 if rng.integers(0, 2) == 0 else A.COMPLEX128
        rdt = np.float32 if prec == A.COMPLEX64 else np.float64
        cdt = np.complex64 if prec == A.COMPLEX64 else np.complex128
        tol = c64_tol(n) if prec == A.COMPLEX64 else C128_TOL
        x = synth.uniform_real(n, (n,), rdt)
        with A.Plan.from_desc(ctx, (n,), precision=prec, kind=A.R2C) as p:
            X = np.empty(n // 2 + 1, cdt)
            p.forward(X, x)
            y = np.empty_like(x)
            p.inverse(y, X)
        assert rel_l2(X, O.rfft(x.astype(np.float64))) <= tol, n
        assert rel_l2(y, x) <= 2 * tol, n


def test_device_memory_stable_over_many_plans(ctx):
    """Creating and destroying hundreds of plans of every kind, clones and
    ragged batches returns device memory to its starting level
    (stress_test.go:71-151, memory stability)."""
    torch.cuda.synchronize()
    shapes = [(4096,), (1 << 16,), (3000,), (8209,), (64, 64), (16, 16, 16), (1 << 20,)]

    def cycle(k):
        s = shapes[k % len(shapes)]
        with A.Plan.from_desc(ctx, s, batch=2) as p:
            x = synth.uniform_complex(k, (2,) + s, np.complex64)
            y = np.empty_like(x)
            p.forward(y, x)
            with p.clone() as c:
                c.forward(y, x)
        if k % 10 == 0:
            with A.Ragged(ctx, [16, 100, 1009, 4096]) as r:
                d = torch.zeros(len(r), dtype=torch.complex64, device="cuda:0")
                e = torch.empty_like(d)
                r.exec_graph(A.FORWARD, e.data_ptr(), d.data_ptr())
                torch.cuda.synchronize()

    for k in range(len(shapes)):       # warm caches and the runtime's pools
        cycle(k)
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for k in range(300):
        cycle(k)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    print(f"free before {free0 >> 20} MiB, after {free1 >> 20} MiB")
    assert free0 - free1 <= 64 << 20
