"""Planner pieces that need no device: candidate sets, features, the shape
generator, and the FP64 cost-model restatement in the oracle (pinned by
central finite differences)."""
import math

import numpy as np

import algofft_hip as A
from algofft_hip import planner
import pyoracle as O


def test_candidate_sets():
    c = planner.candidates({"n": 4096, "batch": 16, "precision": A.COMPLEX64, "stride": 1})
    assert (A.STRATEGY_STOCKHAM, 0) in c and (A.STRATEGY_MIXED, 0) in c
    assert any(s == A.STRATEGY_SIXSTEP for s, _ in c)
    c = planner.candidates({"n": 1 << 20, "batch": 1, "precision": A.COMPLEX64, "stride": 1})
    assert all(s == A.STRATEGY_SIXSTEP for s, _ in c) and len(c) >= 3
    c = planner.candidates({"n": 3000, "batch": 1, "precision": A.COMPLEX128, "stride": 1})
    assert (A.STRATEGY_MIXED, 0) in c and (A.STRATEGY_BLUESTEIN, 0) in c
    c = planner.candidates({"n": 1009, "batch": 1, "precision": A.COMPLEX64, "stride": 1})
    assert c == [(A.STRATEGY_BLUESTEIN, 0)]


def test_features_are_finite_and_distinguish_candidates():
    sh = {"n": 65536, "batch": 8, "precision": A.COMPLEX64, "stride": 1}
    fs = [planner.features(sh, s, p) for s, p in planner.candidates(sh)]
    assert all(np.all(np.isfinite(f)) and f.shape == (32,) for f in fs)
    assert len({f.tobytes() for f in fs}) == len(fs)


def test_shape_generator_matches_the_workload():
    shapes = planner.tuner_shapes(2024, 3000)
    assert len(shapes) == 3000
    ns = np.array([s["n"] for s in shapes])
    assert ns.min() >= 16 and ns.max() <= (1 << 24)
    assert any(s["stride"] > 1 for s in shapes)
    assert {s["precision"] for s in shapes} == {A.COMPLEX64, A.COMPLEX128}
    for s in shapes:
        cap = (1 << 24) if s["precision"] == A.COMPLEX64 else (1 << 23)
        assert s["n"] * s["batch"] <= cap and s["n"] * s["stride"] <= max(cap, s["n"])


def test_oracle_mlp_gradient_by_finite_differences():
    rng = np.random.default_rng(0)
    npar = 32 * 64 + 64 + 64 * 64 + 64 + 64 + 1
    P = rng.standard_normal(npar) * 0.2
    X = rng.standard_normal((6, 32))
    T = rng.standard_normal(6)
    loss, G = O.mlp_loss_grad(P, X, T)
    h = 1e-6
    for idx in rng.choice(npar, 40, replace=False).tolist() + [npar - 1, 0, 32 * 64 + 3]:
        Pp, Pm = P.copy(), P.copy()
        Pp[idx] += h
        Pm[idx] -= h
        fd = (O.mlp_loss_grad(Pp, X, T)[0] - O.mlp_loss_grad(Pm, X, T)[0]) / (2 * h)
        assert abs(fd - G[idx]) <= 1e-6 + 1e-5 * abs(fd), (idx, fd, G[idx])
    assert math.isfinite(loss)


def test_default_smooth_fourstep_split_is_a_candidate():
    # the built-in routing's split (plan.cpp gen_fourstep_split, mirrored by
    # planner.fourstep_split) is among the candidates the tuner and the cost
    # model rank; picks checked against the timed sweep it was fitted on
    want = {(A.COMPLEX64, 18000): 125, (A.COMPLEX64, 24000): 120, (A.COMPLEX64, 100000): 200,
            (A.COMPLEX64, 1000000): 1000, (A.COMPLEX128, 9000): 75, (A.COMPLEX128, 36000): 180}
    for (prec, n), s in want.items():
        assert planner.fourstep_split(n, prec) == s, (n, prec)
    for prec in (A.COMPLEX64, A.COMPLEX128):
        for n in (18000, 20000, 36000, 72000, 100000, 250000, 480000, 1000000, 344064):
            s = planner.fourstep_split(n, prec)
            assert n % s == 0
            c = planner.candidates({"n": n, "batch": 4, "precision": prec, "stride": 1})
            assert (A.STRATEGY_SIXSTEP, s) in c, (n, prec, s, c)
