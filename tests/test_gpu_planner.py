"""Learned planning on the GPU: the cost-model gradients against the FP64
oracle restatement, training, wisdom in the reference's text format
(internal/planner/wisdom.go:96-205), plans that follow wisdom or the model,
and the model's regret against exhaustive timing on held-out shapes."""
import ctypes
import os

import numpy as np
import pytest

import algofft_hip as A
from algofft_hip import planner, synth
import pyoracle as O
from helpers import c64_tol, rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RECORDS = os.path.join(ROOT, "algo-fft_amd", "data", "tune_records_gfx950.csv")


def _identity_norm(m):
    z = np.zeros(32, np.float32)
    o = np.ones(32, np.float32)
    A.check(A.lib().afft_mlp_set_normalization(m.h, m._fp(z), m._fp(o), 0.0, 1.0))


def test_gpu_gradient_matches_oracle(ctx):
    m = planner.CostModel(ctx, seed=3)
    _identity_norm(m)
    rng = np.random.default_rng(1)
    X = rng.standard_normal((300, 32)).astype(np.float32)
    y = rng.standard_normal(300).astype(np.float32)
    loss, g = m.loss_grad(X, y)
    loss_o, g_o = O.mlp_loss_grad(m.params().astype(np.float64), X.astype(np.float64), y.astype(np.float64))
    assert abs(loss - loss_o) <= 1e-4 * abs(loss_o)
    assert np.linalg.norm(g - g_o) / np.linalg.norm(g_o) <= 1e-4
    m.close()


def test_training_fits_a_smooth_function(ctx):
    rng = np.random.default_rng(2)
    X = rng.standard_normal((2000, 32)).astype(np.float32)
    y = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] + 0.1 * X[:, 3:].sum(1)).astype(np.float32)
    m = planner.CostModel(ctx, seed=4)
    m.set_normalization(X, y)
    l0, _ = m.loss_grad(X, y)
    l1 = m.fit(X, y, epochs=60, batch=128, lr=3e-3)
    print("loss", l0, "->", l1)
    assert l1 < 0.1 * l0
    m.close()


def test_weights_file_roundtrip_and_checksum(ctx, tmp_path):
    m = planner.CostModel(ctx, seed=5)
    p = str(tmp_path / "model.bin")
    m.save(p)
    m2 = planner.CostModel(ctx, seed=6)
    m2.load(p)
    assert np.array_equal(m.params(), m2.params())
    raw = bytearray(open(p, "rb").read())
    raw[100] ^= 1
    open(p, "wb").write(bytes(raw))
    with pytest.raises(A.ErrInvalidArgument):
        m2.load(p)
    m.close()
    m2.close()


def _export(ctx):
    need = ctypes.c_int64()
    A.lib().afft_wisdom_export(ctx.h, None, 0, ctypes.byref(need))
    buf = ctypes.create_string_buffer(need.value)
    A.check(A.lib().afft_wisdom_export(ctx.h, buf, need.value, ctypes.byref(need)))
    return buf.value.decode()


def test_wisdom_tune_export_import_and_use():
    ctx = A.Context(0)
    shapes = [{"n": 1 << 18, "batch": 4, "precision": A.COMPLEX64, "stride": 1},
              {"n": 3000, "batch": 32, "precision": A.COMPLEX64, "stride": 1}]
    for sh in shapes:
        d = planner.desc_for(sh)
        cands = planner.candidates(sh)
        times = (ctypes.c_double * len(cands))()
        bs, bp, bm = ctypes.c_int(), ctypes.c_int64(), ctypes.c_double()
        A.check(A.lib().afft_tune_plan(ctx.h, ctypes.byref(d), 1, 3, ctypes.byref(bs), ctypes.byref(bp),
                                       ctypes.byref(bm), times, len(cands)))
        assert bm.value > 0 and all(t > 0 for t in times)
    text = _export(ctx)
    lines = text.strip().split("\n")
    assert len(lines) == 2
    # features: the device word and the shape class (batch 4: bucket 1, batch 32: bucket 2)
    want_feat = {1 << 18: 0x950 | (1 << 16), 3000: 0x950 | (2 << 16)}
    for sh in shapes:
        assert A.lib().afft_wisdom_features(ctx.h, ctypes.byref(planner.desc_for(sh))) == want_feat[sh["n"]]
    for ln in lines:
        size, prec, feat, algo, ts = ln.split(":")
        assert int(size) in (1 << 18, 3000) and prec == "0" and int(feat) == want_feat[int(size)] and int(ts) > 0
        assert algo.startswith("hip_")
    assert int(lines[0].split(":")[0]) < int(lines[1].split(":")[0])   # sorted by size
    ctx.close()
    # a fresh context follows imported wisdom
    ctx2 = A.Context(0)
    A.check(A.lib().afft_wisdom_import(ctx2.h, text.encode()))
    assert _export(ctx2) == text
    with pytest.raises(A.ErrInvalidArgument):
        A.check(A.lib().afft_wisdom_import(ctx2.h, b"12:0:2384:hip_bogus:5\n"))
    forced = "hip_sixstep_256"
    A.check(A.lib().afft_wisdom_import(ctx2.h, f"{1 << 18}:0:{0x950 | (1 << 16)}:{forced}:1\n".encode()))
    x = synth.uniform_complex(1, (4, 1 << 18), np.complex64)
    with A.Plan.from_desc(ctx2, (1 << 18,), batch=4) as p:
        assert p.describe() == "hip_fourstep256x1024", p.describe()
        y = np.empty_like(x)
        p.forward(y, x)
    assert rel_l2(y, O.fft_batch(x.astype(np.complex128))) <= c64_tol(1 << 18)
    # the entry was timed for batches 2-15: a batch of 64 keeps the built-in route
    with A.Plan.from_desc(ctx2, (1 << 18,), batch=64) as p:
        assert p.describe() != "hip_fourstep256x1024", p.describe()
    ctx2.close()


def _split(rows):
    g = planner.group(rows)
    keys = sorted(g)
    return g, keys[:-500], keys[-500:]


@pytest.mark.skipif(not os.path.exists(RECORDS), reason="tuner records not collected")
def test_cost_model_regret_on_held_out_shapes(ctx):
    g, train, test = _split(planner.load_records(RECORDS))
    X, y = planner.dataset(g, train)
    m = planner.CostModel(ctx, seed=11)
    loss = m.fit(X, y, epochs=200, batch=256, lr=2e-3, seed=3)
    r_model, r_model_max = planner.regret(g, test, planner.model_pick(m))
    r_heur, r_heur_max = planner.regret(g, test, planner.heuristic_pick)
    multi = sum(1 for k in test if len(g[k][1]) > 1)
    print(f"train rows {len(y)}, loss {loss:.4f}; held-out shapes {len(test)} ({multi} with a choice): "
          f"model regret mean {r_model:.4f} max {r_model_max:.3f}; heuristic mean {r_heur:.4f} max {r_heur_max:.3f}")
    assert r_model <= 0.10
    # what the network gains over the method without it (the built-in routing)
    assert r_model <= r_heur
    m.close()


@pytest.mark.skipif(not os.path.exists(RECORDS), reason="tuner records not collected")
def test_model_attached_to_context_plans_correctly():
    ctx = A.Context(0)
    g, train, _ = _split(planner.load_records(RECORDS))
    X, y = planner.dataset(g, train)
    m = planner.CostModel(ctx, seed=11)
    m.fit(X, y, epochs=50, batch=256, lr=2e-3, seed=3)
    m.attach(ctx)
    for n in (1 << 20, 12000, 2048):
        sh = {"n": n, "batch": 2, "precision": A.COMPLEX64, "stride": 1}
        cands = planner.candidates(sh)
        pred = m.predict(np.array([planner.features(sh, s, p) for s, p in cands]))
        want = cands[int(np.argmin(pred))]
        x = synth.uniform_complex(n, (2, n), np.complex64)
        with A.Plan.from_desc(ctx, (n,), batch=2) as p:
            algo = p.describe()
            out = np.empty_like(x)
            p.forward(out, x)
        print(n, want, algo)
        if want[0] == A.STRATEGY_SIXSTEP:
            # fourstep<n1>x<n2> (powers of two) or mixed4step<n1>x<n2>
            assert ("fourstep" in algo or "mixed4step" in algo) and f"step{want[1]}x" in algo
        elif want[0] == A.STRATEGY_MIXED:
            assert "mixed" in algo
        elif want[0] == A.STRATEGY_STOCKHAM:
            assert algo.startswith("hip_lds")
        assert rel_l2(out, O.fft_batch(x.astype(np.complex128))) <= c64_tol(n)
    A.check(A.lib().afft_context_set_cost_model(ctx.h, None))
    m.close()
    ctx.close()


@pytest.mark.skipif(not os.path.exists(RECORDS), reason="tuner records not collected")
def test_cost_model_regret_on_freshly_timed_shapes(ctx):
    """Regret on shapes the records never saw, timed now by the autotuner
    (afft_tune_plan, every candidate): 300 seeded shapes from the planner
    workload's distribution with a seed the records do not use.  The network
    is trained on all stored records with three seeds; each must stay within
    10 % mean regret, and their mean must not lose to the built-in routing
    (the method without the network) on the same fresh timings."""
    g_rec = planner.group(planner.load_records(RECORDS))
    X, y = planner.dataset(g_rec, sorted(g_rec))
    shapes = planner.tuner_shapes(9001, 300)
    rows = planner.collect(ctx, shapes, 1, 3)
    g = planner.group(rows)
    keys = [k for k in sorted(g) if len(g[k][1]) > 1]
    assert len(keys) >= 200
    r_heur, r_heur_max = planner.regret(g, keys, planner.heuristic_pick)
    res = []
    for seed in (11, 12, 13):
        m = planner.CostModel(ctx, seed=seed)
        m.fit(X, y, epochs=200, batch=256, lr=2e-3, seed=seed)
        res.append(planner.regret(g, keys, planner.model_pick(m)))
        m.close()
    mean_model = float(np.mean([r[0] for r in res]))
    print(f"fresh shapes with a choice: {len(keys)}; built-in routing regret mean {r_heur:.4f} max {r_heur_max:.3f}; "
          f"network regret per seed " + ", ".join(f"{r[0]:.4f} (max {r[1]:.3f})" for r in res) +
          f"; mean {mean_model:.4f}")
    assert all(r[0] <= 0.10 for r in res)
    assert mean_model <= r_heur
