"""Learned planning on top of the C ABI: the tuner's shape set, timing-record
collection (afft_tune_plan), the MLP cost model (afft_mlp_*, trained by the
library's HIP kernels) and its regret against exhaustive timing.

    python -m algofft_hip.planner collect --count 3000 --out records.csv

Shape set (the planner workload): N log-uniform in 2^4..2^24 (about half
powers of two, a third 2/3/5/7-smooth, the rest primes, primes capped at
2^18 so their Bluestein tables stay quick to build), batch log-uniform in
1..2^16, unit stride or a column stride in 256..4096, complex64 or
complex128; each shape capped at 2^24 (complex64) / 2^23 (complex128)
elements.
"""
import argparse
import csv
import ctypes
import math
import sys
import time

import numpy as np

from . import (COMPLEX64, COMPLEX128, STRATEGY_BLUESTEIN, STRATEGY_MIXED, STRATEGY_SIXSTEP, STRATEGY_STOCKHAM,
               Context, check, lib)
from ._capi import PlanDesc
from . import synth

NFEAT = 32


def tuner_shapes(seed, count):
    u = synth.u01(seed, 6 * count).reshape(count, 6)
    shapes = []
    for i in range(count):
        prec = COMPLEX64 if u[i, 4] < 0.7 else COMPLEX128
        cap = (1 << 24) if prec == COMPLEX64 else (1 << 23)
        kind = u[i, 1]
        hi = 24 if kind < 0.9 else 18
        v = int(round(2 ** (4 + (hi - 4) * u[i, 0])))
        if kind < 0.5:
            n = 1 << int(round(math.log2(v)))
        elif kind < 0.9:
            n = v
            d = 0
            while True:
                if synth._is_smooth7(v - d) and v - d >= 16:
                    n = v - d
                    break
                if synth._is_smooth7(v + d):
                    n = v + d
                    break
                d += 1
        else:
            n = v
            d = 0
            while True:
                if synth._is_prime(v - d) and v - d >= 17:
                    n = v - d
                    break
                if synth._is_prime(v + d):
                    n = v + d
                    break
                d += 1
        n = min(n, cap)
        batch = int(round(2 ** (16 * u[i, 2])))
        stride = 1
        if u[i, 3] < 0.25:
            stride = int(round(2 ** (8 + 4 * u[i, 5])))
            batch = min(batch, stride)
            if n * stride > cap:
                stride = 1
        batch = max(1, min(batch, cap // n))
        shapes.append({"n": n, "batch": batch, "precision": prec, "stride": stride})
    return shapes


def desc_for(shape, strategy=0, split=0):
    d = PlanDesc()
    d.kind = 0
    d.precision = shape["precision"]
    d.rank = 1
    d.dims[0] = shape["n"]
    d.batch = shape["batch"]
    s = shape["stride"]
    if s > 1:
        d.istride = d.ostride = s
        d.idist = d.odist = 1
    d.strategy = strategy
    d.split = split
    return d


def candidates(shape):
    d = desc_for(shape)
    k = lib().afft_plan_candidates(ctypes.byref(d), None, None, 0)
    st = (ctypes.c_int * max(k, 1))()
    sp = (ctypes.c_int64 * max(k, 1))()
    lib().afft_plan_candidates(ctypes.byref(d), st, sp, k)
    return [(st[i], sp[i]) for i in range(k)]


def features(shape, strategy, split):
    d = desc_for(shape)
    out = (ctypes.c_float * NFEAT)()
    check(lib().afft_plan_features(ctypes.byref(d), strategy, split, out))
    return np.frombuffer(out, dtype=np.float32).copy()


# plan.cpp fourstep_split: log2 of the column length where the tuning records
# beat the balanced split, keyed by (log2 N, precision)
TUNED_SPLIT = {(20, COMPLEX64): 9, (23, COMPLEX64): 12, (24, COMPLEX64): 13,
               (18, COMPLEX128): 10, (19, COMPLEX128): 8, (23, COMPLEX128): 12}


def _smooth_radices(m):
    """plan.cpp smooth_schedule: odd radices first, then the power-of-two
    remainder radix and radix-16 passes (powers of two: remainder first)."""
    a = 0
    while m % 2 == 0:
        m //= 2
        a += 1
    odd = []
    for q in (3, 5, 7):
        while m % q == 0:
            m //= q
            odd.append(q)
    if m != 1:
        return []
    tail = ([1 << (a % 4)] if a % 4 else []) + [16] * (a // 4)
    return odd + tail


def fourstep_split(n, prec):
    """plan.cpp gen_fourstep_split: the default n1 of a smooth four-step."""
    maxp = 16384 if prec == COMPLEX64 else 8192
    gcap = 8192 if prec == COMPLEX64 else 4096
    rcap = min(gcap, 4096)

    def pow2(m):
        return m & (m - 1) == 0

    def usable(m):
        return m <= maxp if pow2(m) else m <= rcap

    def cost(m):
        if pow2(m):
            return 3.0
        r = _smooth_radices(m)
        return len(r) + r.count(2) + 0.5 * r.count(4)

    rt = math.sqrt(n)
    n1, best, bal = 0, 0.0, 0.0
    for d in range(max(8, math.ceil(0.2 * rt)), min(n // 8, int(5.0 * rt)) + 1):
        if n % d == 0 and usable(d) and usable(n // d):
            c, b = cost(d) + cost(n // d), abs(math.log(d / rt))
            if not n1 or c < best - 1e-9 or (c < best + 1e-9 and b < bal - 1e-9):
                n1, best, bal = d, c, b
    if not n1:
        q = 1
        while q * q <= n:
            if n % q == 0 and q <= gcap and n // q <= gcap:
                n1 = q
            q += 1
    return n1


def heuristic_choice(shape, cands):
    """The built-in routing (plan.cpp emit_axis) expressed as a candidate."""
    n, prec = shape["n"], shape["precision"]
    maxp = 16384 if prec == COMPLEX64 else 8192
    gcap = 8192 if prec == COMPLEX64 else 4096
    # contiguous rows hold twice that in one pass (in place in LDS)
    rcap = 2 * gcap if shape.get("stride", 1) <= 1 else gcap
    pow2 = n & (n - 1) == 0
    smooth = synth._is_smooth7(n)
    if pow2 and n <= maxp:
        want = (STRATEGY_STOCKHAM, 0)
    elif pow2:
        lg = int(math.log2(n))
        want = (STRATEGY_SIXSTEP, 1 << TUNED_SPLIT.get((lg, prec), lg // 2))
    elif smooth and n <= rcap:
        want = (STRATEGY_MIXED, 0)
    elif smooth:
        want = (STRATEGY_SIXSTEP, fourstep_split(n, prec))
    else:
        want = (STRATEGY_BLUESTEIN, 0)
    return want if want in cands else cands[0]


def collect(ctx, shapes, warmup=1, iters=3, log=None, out_path=None):
    """Times every candidate of every shape; returns rows
    (shape index, n, batch, precision, stride, strategy, split, ms).  With
    out_path the rows are also appended to that CSV as they are measured."""
    rows = []
    t0 = time.time()
    fout = None
    if out_path:
        fout = open(out_path, "w", newline="")
        wr = csv.writer(fout)
        wr.writerow(["shape", "n", "batch", "precision", "stride", "strategy", "split", "ms"])
    for i, sh in enumerate(shapes):
        cands = candidates(sh)
        if not cands:
            continue
        d = desc_for(sh)
        times = (ctypes.c_double * len(cands))()
        bs, bp, bm = ctypes.c_int(), ctypes.c_int64(), ctypes.c_double()
        rc = lib().afft_tune_plan(ctx.h, ctypes.byref(d), warmup, iters, ctypes.byref(bs), ctypes.byref(bp),
                                  ctypes.byref(bm), times, len(cands))
        if rc != 0:
            continue
        for (st, sp), t in zip(cands, times):
            if t > 0:
                rows.append((i, sh["n"], sh["batch"], sh["precision"], sh["stride"], st, sp, t))
                if fout:
                    wr.writerow([i, sh["n"], sh["batch"], sh["precision"], sh["stride"], st, sp, f"{t:.6g}"])
        if fout:
            fout.flush()
        if log and ((i + 1) % 10 == 0 or i < 10):
            print(f"tuned {i + 1}/{len(shapes)} shapes, {len(rows)} records, {time.time() - t0:.0f}s", file=log,
                  flush=True)
    if fout:
        fout.close()
    return rows


def save_records(rows, path):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["shape", "n", "batch", "precision", "stride", "strategy", "split", "ms"])
        for r in rows:
            w.writerow(list(r[:7]) + [f"{r[7]:.6g}"])


def load_records(path):
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append((int(r["shape"]), int(r["n"]), int(r["batch"]), int(r["precision"]), int(r["stride"]),
                         int(r["strategy"]), int(r["split"]), float(r["ms"])))
    return rows


def group(rows):
    """shape index -> (shape dict, [(strategy, split, ms), ...])"""
    g = {}
    for (i, n, b, p, s, st, sp, ms) in rows:
        if i not in g:
            g[i] = ({"n": n, "batch": b, "precision": p, "stride": s}, [])
        g[i][1].append((st, sp, ms))
    return g


class CostModel:
    """The MLP cost model (afft_mlp): predicts log(ms) of (shape, candidate)."""

    def __init__(self, ctx, seed=1):
        h = ctypes.c_void_p()
        check(lib().afft_mlp_create(ctx.h, seed, ctypes.byref(h)))
        self.h = h

    @staticmethod
    def _fp(a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

    def set_normalization(self, X, y):
        mean = X.mean(0).astype(np.float32)
        std = X.std(0).astype(np.float32)
        std[std < 1e-6] = 1.0
        check(lib().afft_mlp_set_normalization(self.h, self._fp(mean), self._fp(std), float(y.mean()),
                                                float(y.std() or 1.0)))

    def fit(self, X, y, epochs=300, batch=256, lr=2e-3, seed=7):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        self.set_normalization(X, y)
        loss = ctypes.c_float()
        check(lib().afft_mlp_train(self.h, self._fp(X), self._fp(y), len(y), epochs, batch, lr, seed,
                                   ctypes.byref(loss)))
        return loss.value

    def predict(self, X):
        X = np.ascontiguousarray(X, np.float32)
        out = np.empty(len(X), np.float32)
        check(lib().afft_mlp_predict(self.h, self._fp(X), len(X), self._fp(out)))
        return out

    def loss_grad(self, X, y):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.empty(lib().afft_mlp_param_count(), np.float32)
        loss = ctypes.c_float()
        check(lib().afft_mlp_loss_grad(self.h, self._fp(X), self._fp(y), len(y), ctypes.byref(loss), self._fp(g)))
        return loss.value, g

    def params(self):
        p = np.empty(lib().afft_mlp_param_count(), np.float32)
        check(lib().afft_mlp_get_params(self.h, self._fp(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        check(lib().afft_mlp_set_params(self.h, self._fp(p)))

    def save(self, path):
        check(lib().afft_mlp_save(self.h, path.encode()))

    def load(self, path):
        check(lib().afft_mlp_load(self.h, path.encode()))

    def attach(self, ctx):
        check(lib().afft_context_set_cost_model(ctx.h, self.h))

    def close(self):
        if self.h:
            check(lib().afft_mlp_destroy(self.h))
            self.h = None


def dataset(groups, keys):
    X, y = [], []
    for k in keys:
        sh, cands = groups[k]
        for st, sp, ms in cands:
            X.append(features(sh, st, sp))
            y.append(math.log(ms))
    return np.array(X, np.float32), np.array(y, np.float32)


def regret(groups, keys, pick):
    """Mean over shapes of t[pick]/t[best] - 1; `pick(shape, cands)` returns
    an index into cands."""
    r = []
    for k in keys:
        sh, cands = groups[k]
        t = [c[2] for c in cands]
        r.append(t[pick(sh, cands)] / min(t) - 1.0)
    return float(np.mean(r)), float(np.max(r))


def model_pick(model):
    def pick(sh, cands):
        X = np.array([features(sh, st, sp) for st, sp, _ in cands], np.float32)
        return int(np.argmin(model.predict(X)))
    return pick


def heuristic_pick(sh, cands):
    want = heuristic_choice(sh, [(st, sp) for st, sp, _ in cands])
    for i, (st, sp, _) in enumerate(cands):
        if (st, sp) == want:
            return i
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--count", type=int, default=3000)
    c.add_argument("--seed", type=int, default=2024)
    c.add_argument("--out", required=True)
    c.add_argument("--iters", type=int, default=3)
    a = ap.parse_args(argv)
    if a.cmd == "collect":
        with Context(0) as ctx:
            shapes = tuner_shapes(a.seed, a.count)
            rows = collect(ctx, shapes, 1, a.iters, log=sys.stdout, out_path=a.out)
        print(f"wrote {len(rows)} records for {len(set(r[0] for r in rows))} shapes to {a.out}")


if __name__ == "__main__":
    main()
