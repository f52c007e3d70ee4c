"""Headline benchmark: batched complex64 C2C forward FFT on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 4096] [--batch 16384]

Workload (the precision configuration's headline point): complex64, N=4096,
batch 16384 in the contiguous layout (transform i at [i*N, (i+1)*N)), 512 MiB
per buffer.  Each row is a sum of 8 complex exponentials at seeded integer
bins with log-uniform amplitudes in [1e-3, 1e1] plus N(0, 1e-3) noise.  A
step is one forward transform of the whole batch, out of place, inputs
already resident in HBM.

Prints one JSON line.  `value` is GFLOP/s with the nominal 5*N*log2(N) flops
per transform; `roofline` prices the dominant kernel against 8 TB/s with the
algorithmic traffic 2*8*N bytes per transform; `cpu_baseline` times the CPU
restatement of the reference's complex64 ForwardBatch on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "algo-fft_amd"))

HBM_PEAK_GBS = 8000.0


def gen_input_device(n, batch, seed, torch):
    """Eight seeded tones per row (amplitudes 1e-3..10, random bins and
    phases), generated on the device in float64 then rounded to complex64."""
    dev = torch.device("cuda:0")
    g = torch.Generator(device="cpu").manual_seed(seed)
    tones = 8
    bins = torch.randint(0, n, (batch, tones), generator=g).to(dev)
    amps = (10.0 ** (-3.0 + 4.0 * torch.rand((batch, tones), generator=g, dtype=torch.float64))).to(dev)
    phase = (2 * np.pi * torch.rand((batch, tones), generator=g, dtype=torch.float64)).to(dev)
    out = torch.empty((batch, n), dtype=torch.complex64, device=dev)
    t = torch.arange(n, device=dev, dtype=torch.int64)
    chunk = max(1, (1 << 22) // n)
    for b0 in range(0, batch, chunk):
        b1 = min(batch, b0 + chunk)
        acc = torch.zeros((b1 - b0, n), dtype=torch.complex128, device=dev)
        for j in range(tones):
            ph = ((bins[b0:b1, j:j + 1] * t[None, :]) % n).to(torch.float64) * (2 * np.pi / n) + phase[b0:b1, j:j + 1]
            acc += amps[b0:b1, j:j + 1] * torch.exp(1j * ph)
        noise = torch.randn((b1 - b0, n, 2), dtype=torch.float64, device=dev) * 1e-3
        acc += torch.complex(noise[..., 0], noise[..., 1])
        out[b0:b1] = acc.to(torch.complex64)
    return out


def traffic_bytes(n, batch):
    """HBM bytes per launch measured by rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
    for this workload, scaled by the same counters on a device copy of known
    size (profiles/traffic_*.json, written from those runs), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_c64_n{n}_b{batch}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("traffic_bytes_per_launch")


def _time_plan(torch, plan, A, stream, y, x, steps, warm_s=0.1):
    """Mean time of `steps` back-to-back executions, after `warm_s` seconds of
    untimed ones: the host work between legs (plan creation, accuracy checks
    on the CPU) idles the GPU long enough for its clocks to drop."""
    sh = stream.cuda_stream
    t0 = time.time()
    while True:
        plan.exec_device(A.FORWARD, y.data_ptr(), x.data_ptr(), sh)
        torch.cuda.synchronize()
        if time.time() - t0 >= warm_s:
            break
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        plan.exec_device(A.FORWARD, y.data_ptr(), x.data_ptr(), sh)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def _accuracy(x, y, rows):
    ref = np.fft.fft(x[rows].cpu().numpy().astype(np.complex128), axis=1)
    got = y[rows].cpu().numpy().astype(np.complex128)
    rel = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    mx = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    return rel, mx


def mfma_leg(torch, A, ctx, stream, x, steps, n, batch):
    """The same headline workload through the MFMA form (FP16 hi/lo radix-16
    products, FP32 accumulation): time and accuracy beside the plain path."""
    plan = A.Plan.from_desc(ctx, (n,), precision=A.COMPLEX64, batch=batch, flags=A.FLAG_MFMA)
    y = torch.empty_like(x)
    ms = _time_plan(torch, plan, A, stream, y, x, steps)
    rel, mx = _accuracy(x, y, np.linspace(0, batch - 1, 16).astype(np.int64))
    algo = plan.describe()
    plan.close()
    return {"algorithm": algo, "ms_per_step": round(ms, 5),
            "gflops": round(5.0 * n * np.log2(n) * batch / (ms * 1e-3) / 1e9, 1), "rel_l2_vs_fp64": rel,
            "max_err_vs_fp64": mx}


def sweep_leg(torch, A, ctx, stream, steps, x4096=None):
    """Precision workload at every size: complex64 N in {256, 4096, 65536,
    2^20}, batch 2^26/N (512 MiB per buffer), plain FP32 and MFMA forms, each
    timed and checked against numpy's FP64 FFT on sampled rows."""
    res = []
    for n in (256, 4096, 65536, 1 << 20):
        batch = (1 << 26) // n
        x = x4096 if (n == 4096 and x4096 is not None) else gen_input_device(n, batch, 2024 + n, torch)
        y = torch.empty_like(x)
        rows = np.linspace(0, batch - 1, 8).astype(np.int64)
        for form, flags in (("plain", 0), ("mfma", A.FLAG_MFMA)):
            plan = A.Plan.from_desc(ctx, (n,), precision=A.COMPLEX64, batch=batch, flags=flags)
            ms = _time_plan(torch, plan, A, stream, y, x, steps)
            rel, mx = _accuracy(x, y, rows)
            gbs = 16.0 * n * batch / (ms * 1e-3) / 1e9
            res.append({"n": n, "batch": batch, "form": form, "algorithm": plan.describe(), "ms": round(ms, 5),
                        "gbs_algorithmic": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4),
                        "gflops": round(5.0 * n * np.log2(n) * batch / (ms * 1e-3) / 1e9, 1),
                        "rel_l2_vs_fp64": rel, "max_err_vs_fp64": mx, "rel_l2_budget": 2e-6 * np.log2(n)})
            plan.close()
        del x, y
        torch.cuda.empty_cache()
    return res


def smooth_leg(torch, A, ctx, stream, steps):
    """2/3/5/7-smooth lengths (the mixed-radix path) at batch 2^26/N, 512 MiB
    per buffer: N = 1000, 3000, 6000 and 12000 in one pass (6000 and 12000
    with every pass in place in one LDS buffer), 100000 as four-step, and the
    2-D 1000 x 1000 (rows, then the strided column axis)."""
    res = []
    g = torch.Generator(device="cuda:0").manual_seed(17)
    for n in (1000, 3000, 6000, 12000, 100000):
        batch = (1 << 26) // n
        x = torch.complex(torch.rand(batch * n, device="cuda:0", generator=g) * 2 - 1,
                          torch.rand(batch * n, device="cuda:0", generator=g) * 2 - 1)
        y = torch.empty_like(x)
        plan = A.Plan.from_desc(ctx, (n,), batch=batch)
        ms = _time_plan(torch, plan, A, stream, y, x, steps)
        rows = np.linspace(0, batch - 1, 4).astype(np.int64)
        rel, mx = _accuracy(x.view(batch, n), y.view(batch, n), rows)
        gbs = 16.0 * n * batch / (ms * 1e-3) / 1e9
        res.append({"n": n, "batch": batch, "algorithm": plan.describe(), "ms": round(ms, 5),
                    "gbs_algorithmic": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4),
                    "rel_l2_vs_fp64": rel, "max_err_vs_fp64": mx})
        plan.close()
        del x, y
    d = 1000
    batch = (1 << 26) // (d * d)
    x = torch.complex(torch.rand(batch * d * d, device="cuda:0", generator=g) * 2 - 1,
                      torch.rand(batch * d * d, device="cuda:0", generator=g) * 2 - 1)
    y = torch.empty_like(x)
    plan = A.Plan.from_desc(ctx, (d, d), batch=batch)
    ms = _time_plan(torch, plan, A, stream, y, x, steps)
    sel = [0, batch - 1]
    ref = np.fft.fft2(x.view(batch, d, d)[sel].cpu().numpy().astype(np.complex128))
    got = y.view(batch, d, d)[sel].cpu().numpy().astype(np.complex128)
    gbs = 16.0 * d * d * batch / (ms * 1e-3) / 1e9
    res.append({"n": [d, d], "batch": batch, "algorithm": plan.describe(), "ms": round(ms, 5),
                "gbs_algorithmic": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4),
                "rel_l2_vs_fp64": float(np.linalg.norm(got - ref) / np.linalg.norm(ref)),
                "max_err_vs_fp64": float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))})
    plan.close()
    del x, y
    torch.cuda.empty_cache()
    return res


def nd_leg(torch, A, ctx, stream, steps):
    """N-D, real and FP64 workloads: float32 R2C 512^3, complex128 C2C
    8192^2 and complex128 Bluestein N=1,000,003 x 8, each forward, device
    resident; algorithmic bytes as sizeof(in) + sizeof(out)."""
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(11)
    res = []

    def run(name, plan, x, y, alg_bytes, flops, check):
        ms = _time_plan(torch, plan, A, stream, y, x, steps)
        gbs = alg_bytes / (ms * 1e-3) / 1e9
        entry = {"workload": name, "algorithm": plan.describe(), "ms": round(ms, 4),
                 "gbs_algorithmic": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4),
                 "gflops": round(flops / (ms * 1e-3) / 1e9, 1)}
        entry.update(check())
        res.append(entry)

    d = 512
    xr = torch.rand((d, d, d), device=dev, generator=g) * 2 - 1
    X = torch.empty((d, d, d // 2 + 1), dtype=torch.complex64, device=dev)
    plan = A.Plan.from_desc(ctx, (d, d, d), kind=A.R2C)

    def chk_r2c():
        # Parseval over the half spectrum (interior bins twice)
        Xd = X.to(torch.complex128)
        e = (Xd.abs() ** 2).sum() * 2 - (Xd[..., 0].abs() ** 2).sum() - (Xd[..., -1].abs() ** 2).sum()
        t = (xr.to(torch.float64) ** 2).sum() * d ** 3
        return {"parseval_rel_dev": abs(float(e / t) - 1.0)}

    run("float32 R2C 3-D 512x512x512", plan, xr, X, 4 * d ** 3 + 8 * d * d * (d // 2 + 1),
        2.5 * d ** 3 * np.log2(float(d ** 3)), chk_r2c)
    plan.close()
    del xr, X
    n = 8192
    x = torch.complex(torch.randn((n, n), device=dev, dtype=torch.float64, generator=g),
                      torch.randn((n, n), device=dev, dtype=torch.float64, generator=g))
    y = torch.empty_like(x)
    plan = A.Plan.from_desc(ctx, (n, n), precision=A.COMPLEX128)

    def chk_2d():
        rel = float((y.abs() ** 2).sum() / ((x.abs() ** 2).sum() * n * n)) - 1.0
        k1 = 77
        colmix = np.exp(-2j * np.pi * k1 * np.arange(n) / n) @ x.cpu().numpy()
        want = np.fft.fft(colmix)
        got = y[k1].cpu().numpy()
        return {"parseval_rel_dev": abs(rel), "sampled_row_rel_l2": float(np.linalg.norm(got - want) / np.linalg.norm(want))}

    run("complex128 C2C 2-D 8192x8192", plan, x, y, 2 * 16 * n * n, 5.0 * n * n * np.log2(float(n * n)), chk_2d)
    plan.close()
    del x, y
    n, b = 1_000_003, 8
    x = torch.complex(torch.randn((b, n), device=dev, dtype=torch.float64, generator=g),
                      torch.randn((b, n), device=dev, dtype=torch.float64, generator=g))
    y = torch.empty_like(x)
    plan = A.Plan.from_desc(ctx, (n,), precision=A.COMPLEX128, batch=b)

    def chk_blue():
        ref = np.fft.fft(x[:2].cpu().numpy(), axis=1)
        got = y[:2].cpu().numpy()
        return {"rel_l2_vs_fp64": float(np.linalg.norm(got - ref) / np.linalg.norm(ref)), "rel_l2_budget": 1e-13}

    run("complex128 Bluestein N=1000003 batch 8", plan, x, y, 2 * 16 * n * b, 5.0 * n * np.log2(n) * b, chk_blue)
    # executed traffic of the chained form per transform, in points of the
    # padded M = 2^21: pass A1 writes M; B1 reads M and the M-point filter
    # and writes M; A2 reads and writes M; B2 reads M (8M), plus the N-point
    # input, chirp, post-chirp and output (4N)
    m = 1 << 21
    res[-1]["executed_bytes"] = (8 * m + 4 * n) * 16 * b
    res[-1]["gbs_executed"] = round(res[-1]["executed_bytes"] / (res[-1]["ms"] * 1e-3) / 1e9, 1)
    plan.close()
    del x, y
    torch.cuda.empty_cache()
    return res


def latency_leg():
    """Host cost per call of a small device-resident transform (complex64
    N=256, one row) from C++: the checked entry point against the fast plan
    (plan_fast.go), enqueue-only and launch-and-wait."""
    import subprocess
    exe = os.path.join(ROOT, "algo-fft_amd", "lib", "latency_bench")
    if not os.path.exists(exe):
        return {"error": "latency_bench not built"}
    r = subprocess.run([exe, "256", "20000"], capture_output=True, text=True, timeout=120)
    try:
        return json.loads(r.stdout.strip().splitlines()[-1])
    except (ValueError, IndexError):
        return {"error": (r.stdout + r.stderr)[-300:]}


def host_leg(torch, A, ctx, n, batch, reps=3):
    """End to end from host memory: H2D, transform, D2H through the host-slice
    entry point (afft_plan_forward), from pageable and from pinned buffers."""
    import ctypes
    plan = A.Plan.from_desc(ctx, (n,), precision=A.COMPLEX64, batch=batch)
    total = n * batch
    out = {"workload": f"complex64 N={n} batch={batch} host in, host out", "bytes_each_way": 8 * total}
    for kind in ("pageable", "pinned"):
        pin = kind == "pinned"
        xs = torch.empty(total, dtype=torch.complex64, pin_memory=pin)
        xs.real.uniform_(-1, 1)
        xs.imag.uniform_(-1, 1)
        ys = torch.empty(total, dtype=torch.complex64, pin_memory=pin)
        fn = A.lib().afft_plan_forward
        args = (plan.h, ctypes.c_void_p(ys.data_ptr()), total, ctypes.c_void_p(xs.data_ptr()), total)
        A.check(fn(*args))
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            A.check(fn(*args))
            ts.append(time.perf_counter() - t0)
        t = float(np.median(ts))
        out[kind] = {"ms": round(t * 1e3, 3), "gflops": round(5.0 * n * np.log2(n) * batch / t / 1e9, 1),
                     "gbs_each_way": round(8 * total / t / 1e9, 2)}
        del xs, ys
    plan.close()
    return out


def ragged_leg(torch, A, ctx, stream, steps):
    """Mixed-size workload: 4096 complex64 transforms, lengths log-uniform in
    [16, 65536] (60% powers of two, 30% 7-smooth, 10% primes via Bluestein),
    packed back to back.  The persistent work-queue launch and its graph
    replay against the plain per-size path: one batched plan per distinct
    length (Plan.ForwardBatch per size) on a copy sorted by length, and one
    plan execution per job.  Accuracy: every job of the graph output against
    numpy's FP64 FFT, as rel L2 over the complex64 budget 2e-6*log2(n)."""
    from algofft_hip import synth
    sizes = synth.ragged_sizes(1234, 4096)
    g = torch.Generator(device="cuda:0").manual_seed(5)
    total = sum(sizes)
    x = torch.complex(torch.rand(total, device="cuda:0", generator=g) * 2 - 1,
                      torch.rand(total, device="cuda:0", generator=g) * 2 - 1)
    y = torch.empty_like(x)
    sh = stream.cuda_stream

    def timed(fn, k):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(k):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / k

    t0 = time.perf_counter()
    rg = A.Ragged(ctx, sizes, A.COMPLEX64)
    create_s = time.perf_counter() - t0
    ms_p = timed(lambda: rg.exec(A.FORWARD, y.data_ptr(), x.data_ptr(), sh), steps)
    ms_g = timed(lambda: rg.exec_graph(A.FORWARD, y.data_ptr(), x.data_ptr(), sh), steps)
    ms_j = timed(lambda: rg.exec_per_job(A.FORWARD, y.data_ptr(), x.data_ptr(), sh), max(2, steps // 10))
    rg.exec_graph(A.FORWARD, y.data_ptr(), x.data_ptr(), sh)
    torch.cuda.synchronize()
    # accuracy of every job of the graph output
    xs, ys = x.cpu().numpy(), y.cpu().numpy()
    ratios, off = [], 0
    for n in sizes:
        ref = np.fft.fft(xs[off:off + n].astype(np.complex128))
        got = ys[off:off + n].astype(np.complex128)
        ratios.append(float(np.linalg.norm(got - ref) / np.linalg.norm(ref)) / (2e-6 * max(1.0, np.log2(n))))
        off += n
    rg.close()
    # the same jobs sorted by length: the per-size comparator's layout (one
    # contiguous batch per distinct length), and the graph on that layout
    order = sorted(range(len(sizes)), key=lambda i: (sizes[i], i))
    ssizes = [sizes[i] for i in order]
    rs = A.Ragged(ctx, ssizes, A.COMPLEX64)
    ms_gs = timed(lambda: rs.exec_graph(A.FORWARD, y.data_ptr(), x.data_ptr(), sh), steps)
    ms_s = timed(lambda: rs.exec_per_size(A.FORWARD, y.data_ptr(), x.data_ptr(), sh), max(3, steps // 2))
    rs.close()
    flops = float(sum(5.0 * n * np.log2(n) for n in sizes))
    alg = 16.0 * total
    del x, y
    torch.cuda.empty_cache()
    return {"workload": "4096 ragged complex64 forward transforms, N log-uniform 16..65536 "
                        "(60% pow2 / 30% 7-smooth / 10% prime)",
            "points": total, "distinct_sizes": len(set(sizes)), "create_s": round(create_s, 3),
            "persistent_ms": round(ms_p, 4), "graph_ms": round(ms_g, 4),
            "graph_sorted_ms": round(ms_gs, 4),
            "per_size_batched_ms": round(ms_s, 4), "per_size_batched_launch_groups": len(set(sizes)),
            "per_job_ms": round(ms_j, 4),
            "gflops_graph": round(flops / (ms_g * 1e-3) / 1e9, 1),
            "gbs_graph_algorithmic": round(alg / (ms_g * 1e-3) / 1e9, 1),
            "speedup_graph_vs_per_size_batched": round(ms_s / ms_g, 2),
            "speedup_graph_vs_per_job": round(ms_j / ms_g, 2),
            "rel_l2_over_budget_max": round(max(ratios), 5),
            "rel_l2_over_budget_mean": round(float(np.mean(ratios)), 5),
            "jobs_checked": len(ratios)}


def plan_create_leg(A, ctx, reps=5):
    """Host time of plan creation (NewPlan; the reference publishes 4.48 ms
    for NewPlan at N=65536, BENCHMARKS.md:129-131): complex64 N=65536,
    complex128 Bluestein N=1,000,003 (filter spectra of the padded 2^21
    points), median of `reps` creations each."""
    out = {}
    for name, dims, prec in (("complex64 N=65536", (65536,), A.COMPLEX64),
                             ("complex64 N=4096 batch 16384", (4096,), A.COMPLEX64),
                             ("complex128 Bluestein N=1000003", (1_000_003,), A.COMPLEX128)):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            p = A.Plan.from_desc(ctx, dims, precision=prec, batch=16384 if dims == (4096,) else 1)
            ts.append(time.perf_counter() - t0)
            p.close()
        out[name] = {"ms_median": round(1e3 * float(np.median(ts)), 3), "ms_min": round(1e3 * min(ts), 3)}
    return out


def cpu_baseline(n, seconds=10.0):
    """CPU restatement of the reference's complex64 ForwardBatch (oracle),
    sharded over host threads; timed on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    from algofft_hip import synth

    # the GPU box allots 16 host CPUs to one GPU, whatever the machine has:
    # use them all, or every CPU this process may run on if that is fewer
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, min(16, avail))
    sample = 2048
    x = synth.uniform_complex(99, (sample, n), np.complex64)
    O.fft_batch(x[:threads], nthreads=threads)  # warm
    reps, t0 = 0, time.perf_counter()
    while True:
        O.fft_batch(x, nthreads=threads)
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    flops = 5.0 * n * np.log2(n) * sample * reps
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": round(flops / el / 1e9, 3), "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "cpus_available": avail, "threads": threads, "cpu_model": model,
            "sample": f"{reps}x ForwardBatch of {sample} complex64 N={n} transforms "
                      f"({el:.1f} s, {threads} threads, oracle routes N={n} to "
                      f"{['copy', 'radix-2 Stockham', 'six-step', 'mixed radix', 'Bluestein'][O.route(n)]})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--batch", type=int, default=0, help="default 2^26/N")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-ragged", action="store_true", help="skip the mixed-size leg")
    ap.add_argument("--no-sweep", action="store_true", help="skip the precision sweep over N")
    ap.add_argument("--no-host", action="store_true", help="skip the host-memory end-to-end leg")
    ap.add_argument("--no-nd", action="store_true", help="skip the N-D / real / FP64 leg")
    ap.add_argument("--no-smooth", action="store_true", help="skip the smooth-length leg")
    ap.add_argument("--no-latency", action="store_true", help="skip the small-transform latency probe")
    ap.add_argument("--no-create", action="store_true", help="skip the plan-creation timing leg")
    ap.add_argument("--prewarm", type=float, default=0.5,
                    help="seconds of untimed steps before the warmup steps, to leave the idle clock state")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py runs on one GPU; --gpus must be 1 "
                                   "(batched transforms shard along the batch with no exchange)"}))
        sys.exit(2)

    import torch
    import algofft_hip as A

    n = args.n
    batch = args.batch or (1 << 26) // n
    torch.cuda.set_device(0)
    x = gen_input_device(n, batch, 1234, torch)
    y = torch.empty_like(x)
    torch.cuda.synchronize()

    ctx = A.Context(0)
    plan = A.Plan.from_desc(ctx, (n,), precision=A.COMPLEX64, batch=batch)
    # a dedicated stream: the launches and the timing events share it
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sh = stream.cuda_stream
    assert sh, "need a non-null stream handle"

    def step():
        plan.exec_device(A.FORWARD, y.data_ptr(), x.data_ptr(), sh)

    # Untimed: an idle GPU needs some 0.1 s of continuous work to reach its
    # steady clocks (the first 50 steps after an idle second run about 8 %
    # slower than the steady state), so steps run for `prewarm` seconds
    # before the warmup steps proper.
    t_end = time.perf_counter() + args.prewarm
    while time.perf_counter() < t_end:
        for _ in range(20):
            step()
        torch.cuda.synchronize()
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # whole-job timing: one event pair around the K steps
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1) / args.steps
    # per-launch durations (same stream): the roofline uses their average
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(2 * args.steps)]
    for i in range(args.steps):
        evs[2 * i].record(stream)
        step()
        evs[2 * i + 1].record(stream)
    torch.cuda.synchronize()
    launch_ms = [evs[2 * i].elapsed_time(evs[2 * i + 1]) for i in range(args.steps)]
    kms = float(np.mean(launch_ms))
    # device copy of the same bytes: the practical streaming ceiling
    cps = []
    for i in range(5):
        c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        c0.record(stream)
        y.copy_(x)
        c1.record(stream)
        torch.cuda.synchronize()
        if i:
            cps.append(c0.elapsed_time(c1))
    copy_gbs = 2 * x.numel() * x.element_size() / (np.mean(cps) * 1e-3) / 1e9
    step()
    torch.cuda.synchronize()

    # accuracy on sampled rows against numpy's FP64 FFT of the same input
    rows = np.linspace(0, batch - 1, 16).astype(np.int64)
    xs = x[rows].cpu().numpy().astype(np.complex128)
    ys = y[rows].cpu().numpy().astype(np.complex128)
    ref = np.fft.fft(xs, axis=1)
    rel = float(np.linalg.norm(ys - ref) / np.linalg.norm(ref))

    flops = 5.0 * n * np.log2(n) * batch
    gflops = flops / (ms * 1e-3) / 1e9
    alg_bytes = 2 * 8 * n * batch
    gbs = alg_bytes / (kms * 1e-3) / 1e9
    nsteps = len(plan.describe().split("+"))
    out = {
        "metric": "batched C2C GFLOP/s + % HBM peak, MFMA/ragged vs plain GPU at rel L2<=2e-6*log2N",
        "value": round(gflops, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (f32 arithmetic)",
        "data": "synthetic (8 seeded tones, log-uniform amplitudes 1e-3..1e1, N(0,1e-3) noise)",
        "config": {"workload": f"complex64 C2C forward N={n} batch={batch} contiguous",
                   "n": n, "batch": batch, "algorithm": plan.describe(), "parallelism": "single GPU"},
        "rel_l2_vs_fp64": rel,
        "rel_l2_budget": 2e-6 * np.log2(n),
        "wall_s_timed": round(wall, 4),
        "prewarm_s": args.prewarm,
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic_bytes(n, batch),
                     "frac_source": "mean of per-launch HIP event pairs on the launch stream",
                     "frac_from_ms_per_step": round(alg_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "kernel_ms": round(kms, 5), "kernel_launches_per_step": nsteps,
                     "algorithmic_bytes_per_launch": alg_bytes,
                     "device_copy_gbs": round(copy_gbs, 1), "frac_of_device_copy": round(gbs / copy_gbs, 4)},
    }
    plan.close()
    out["mfma"] = mfma_leg(torch, A, ctx, stream, x, args.steps, n, batch)
    out["mfma"]["speed_vs_plain"] = round(ms / out["mfma"]["ms_per_step"], 3)
    if not args.no_ragged:
        out["ragged"] = ragged_leg(torch, A, ctx, stream, max(3, args.steps // 5))
    if not args.no_sweep:
        out["precision_sweep"] = sweep_leg(torch, A, ctx, stream, max(5, args.steps // 5),
                                           x if (n == 4096 and batch == (1 << 26) // 4096) else None)
    del y
    torch.cuda.empty_cache()
    if not args.no_smooth:
        out["smooth_sizes"] = smooth_leg(torch, A, ctx, stream, max(5, args.steps // 5))
    if not args.no_nd:
        out["nd_real_fp64"] = nd_leg(torch, A, ctx, stream, max(3, args.steps // 10))
    if not args.no_latency:
        out["latency_small_n"] = latency_leg()
    if not args.no_host:
        out["host_end_to_end"] = host_leg(torch, A, ctx, n, batch)
    if not args.no_create:
        out["plan_create"] = plan_create_leg(A, ctx)
    ctx.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
