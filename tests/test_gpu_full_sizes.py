"""The full-size workloads: the precision sweep (complex64 N in {256, 4096,
65536, 2^20}, batch 2^26/N, 8 seeded tones + noise, plain FP32 and MFMA
forms), float32 R2C 512^3,
complex128 C2C 8192^2 and complex128 Bluestein N=1,000,003 x 8.

At full size the oracle checks sampled rows (C2C sweep, Bluestein in full);
the 3-D and 2-D transforms are checked through size-independent properties
(round trip, Parseval, separability along sampled lines, direct DFT sums
at sampled bins) plus the oracle on a reduced shape.
"""
import numpy as np
import pytest
import torch

import algofft_hip as A
from algofft_hip import synth
import pyoracle as O
from helpers import C128_TOL, c64_tol, rel_l2

pytestmark = pytest.mark.gpu


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")


def _exec(plan, direction, out, x, stream):
    plan.exec_device(direction, out.data_ptr(), x.data_ptr(), stream.cuda_stream)
    stream.synchronize()


@pytest.mark.parametrize("form", ["plain", "mfma"])
@pytest.mark.parametrize("n", [256, 4096, 65536, 1 << 20])
def test_precision_sweep_full_batch(ctx, n, form):
    batch = (1 << 26) // n
    rows = 4 if n >= 65536 else 16
    # full-size device input: the seeded tones for sampled rows are checked
    # exactly; the rest of the buffer is filled with the same generator
    x, bins, amps = synth.sinusoids(2024 + n, n, rows)
    xt = torch.empty((batch, n), dtype=torch.complex64, device="cuda:0")
    xt.real.uniform_(-1, 1)
    xt.imag.uniform_(-1, 1)
    pick = np.linspace(0, batch - 1, rows).astype(np.int64)
    xt[torch.from_numpy(pick).cuda()] = _dev(x)
    yt = torch.empty_like(xt)
    s = torch.cuda.Stream()
    flags = A.FLAG_MFMA if form == "mfma" else 0
    with A.Plan.from_desc(ctx, (n,), batch=batch, flags=flags) as p:
        _exec(p, A.FORWARD, yt, xt, s)
        algo = p.describe()
    y = yt[torch.from_numpy(pick).cuda()].cpu().numpy()
    ref = O.fft_batch(x.astype(np.complex128))
    err = rel_l2(y, ref)
    maxerr = float(np.max(np.abs(y - ref)) / np.max(np.abs(ref)))
    print(f"{form} n={n} batch={batch} algo={algo} relL2={err:.3e} max={maxerr:.3e} budget={c64_tol(n):.1e}")
    assert err <= c64_tol(n)


def test_r2c_3d_512_cubed(ctx):
    d = 512
    xr = torch.rand((d, d, d), device="cuda:0", dtype=torch.float32, generator=torch.Generator("cuda").manual_seed(3))
    xr = xr * 2 - 1
    X = torch.empty((d, d, d // 2 + 1), dtype=torch.complex64, device="cuda:0")
    back = torch.empty_like(xr)
    s = torch.cuda.Stream()
    with A.Plan.from_desc(ctx, (d, d, d), kind=A.R2C) as p:
        _exec(p, A.FORWARD, X, xr, s)
        _exec(p, A.INVERSE, back, X, s)
        algo = p.describe()
    print("algo", algo)
    # round trip
    assert (torch.linalg.norm(back - xr) / torch.linalg.norm(xr)).item() <= 2 * c64_tol(d ** 3)
    # Parseval with the half spectrum: interior bins counted twice
    Xd = X.to(torch.complex128)
    e_full = (Xd.abs() ** 2).sum() * 2 - (Xd[..., 0].abs() ** 2).sum() - (Xd[..., -1].abs() ** 2).sum()
    e_time = (xr.to(torch.float64) ** 2).sum() * d ** 3
    assert abs((e_full / e_time).item() - 1.0) < 1e-5
    # sampled bins against a direct DFT sum over all 2^27 inputs
    xn = xr.cpu().numpy().astype(np.float64)
    idx = np.arange(d)
    for (k0, k1, k2) in [(0, 0, 0), (3, 511, 17), (256, 100, 256), (7, 9, 200)]:
        e0 = np.exp(-2j * np.pi * k0 * idx / d)
        e1 = np.exp(-2j * np.pi * k1 * idx / d)
        e2 = np.exp(-2j * np.pi * k2 * idx / d)
        want = np.einsum("i,j,k,ijk->", e0, e1, e2, xn, optimize=True)
        got = X[k0, k1, k2].item()
        assert abs(got - want) <= 1e-6 * np.sqrt(d ** 3) * 10, (k0, k1, k2, got, want)
    # reduced shape against the oracle (PlanReal3D restatement)
    small = synth.uniform_real(5, (32, 16, 64), np.float32)
    with A.Plan.from_desc(ctx, small.shape, kind=A.R2C) as p:
        out = np.empty((32, 16, 33), np.complex64)
        p.forward(out, small)
    assert rel_l2(out, O.rfftn(small.astype(np.float64))) <= c64_tol(small.size)


def test_c128_2d_8192_squared(ctx):
    n = 8192
    g = torch.Generator("cuda").manual_seed(9)
    x = torch.complex(torch.randn((n, n), device="cuda:0", dtype=torch.float64, generator=g),
                      torch.randn((n, n), device="cuda:0", dtype=torch.float64, generator=g))
    X = torch.empty_like(x)
    back = torch.empty_like(x)
    s = torch.cuda.Stream()
    with A.Plan.from_desc(ctx, (n, n), precision=A.COMPLEX128) as p:
        _exec(p, A.FORWARD, X, x, s)
        _exec(p, A.INVERSE, back, X, s)
        print("algo", p.describe())
    assert (torch.linalg.norm(back - x) / torch.linalg.norm(x)).item() <= 4 * C128_TOL
    ratio = ((X.abs() ** 2).sum() / ((x.abs() ** 2).sum() * n * n)).item()
    assert abs(ratio - 1.0) < 1e-12
    # separability: row k of the spectrum equals the FFT over rows of the
    # column-transformed data; check sampled output rows against the oracle
    xc = x.cpu().numpy()
    for k1 in (0, 1234):
        colmix = np.exp(-2j * np.pi * k1 * np.arange(n) / n) @ xc     # DFT along axis 0 at bin k1
        want = O.fft(colmix)
        got = X[k1].cpu().numpy()
        assert rel_l2(got, want) <= 1e-12


def test_c128_bluestein_million(ctx):
    n, batch = 1_000_003, 8
    x = synth.normal_complex(31, (batch, n), np.complex128)
    with A.Plan.from_desc(ctx, (n,), precision=A.COMPLEX128, batch=batch) as p:
        y = np.empty_like(x)
        p.forward(y, x)
        algo = p.describe()
        back = np.empty_like(x)
        p.inverse(back, y)
    ref = O.fft_batch(x[:2])
    err = rel_l2(y[:2], ref)
    print(f"algo={algo} relL2 vs exact-chirp FP64 oracle {err:.3e}")
    assert err <= C128_TOL
    assert rel_l2(back, x) <= 2 * C128_TOL
    # the reference's unreduced chirp (bluestein.go:23-28) moves the result
    # by ~1e-10: reported, not matched
    ref_chirp = O.fft_batch(x[:1], exact_chirp=False)
    print(f"reference-chirp deviation {rel_l2(ref_chirp, ref[:1]):.3e}")
