"""ctypes view of the C oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

Imported by tests/, by __graft_entry__.smoke() and by bench.py's CPU-baseline
leg, and by nothing else.  Every function returns numpy arrays of
complex64/complex128 (or float32/float64 for real signals).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_PREC = {np.complex64: ("c64", np.float32), np.complex128: ("c128", np.float64)}


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(path)
        P, l, i = ctypes.c_void_p, ctypes.c_long, ctypes.c_int
        for sfx in ("c64", "c128"):
            getattr(L, "or_fft_" + sfx).argtypes = [l, P, P, i, i]
            getattr(L, "or_fft_batch_" + sfx).argtypes = [l, l, P, P, i, i, i]
            getattr(L, "or_fft_strided_" + sfx).argtypes = [l, l, P, P, i]
            getattr(L, "or_rfft_" + sfx).argtypes = [l, l, P, P]
            getattr(L, "or_irfft_" + sfx).argtypes = [l, l, P, P]
            getattr(L, "or_fftnd_" + sfx).argtypes = [i, P, P, P, i]
            getattr(L, "or_rfftnd_" + sfx).argtypes = [i, P, P, P]
            getattr(L, "or_convolve_" + sfx).argtypes = [l, P, l, P, P]
            getattr(L, "or_convolve_real_" + sfx).argtypes = [l, P, l, P, P]
            getattr(L, "or_twiddles_" + sfx).argtypes = [l, P]
            getattr(L, "or_chirp_" + sfx).argtypes = [l, i, P]
            getattr(L, "or_route_" + sfx).argtypes = [l]
        L.or_dft_naive.argtypes = [l, P, P, i]
        L.or_rng_u01.argtypes = [ctypes.c_uint64, ctypes.c_uint64, l, P]
        L.or_is_highly_composite.argtypes = [l]
        L.or_next_pow2.argtypes = [l]
        L.or_next_pow2.restype = l
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _sfx(dtype):
    return _PREC[np.dtype(dtype).type][0]


def fft(x, inverse=False, exact_chirp=True):
    """Plan.Forward / Plan.Inverse of one transform (plan.go:332-419)."""
    x = np.ascontiguousarray(x)
    out = np.empty_like(x)
    rc = getattr(lib(), "or_fft_" + _sfx(x.dtype))(x.size, _p(x), _p(out), int(inverse), int(exact_chirp))
    assert rc == 0
    return out


def fft_batch(x, inverse=False, nthreads=1, exact_chirp=True):
    """Plan.ForwardBatch over the rows of a 2-D array (plan.go:504-571)."""
    x = np.ascontiguousarray(x)
    out = np.empty_like(x)
    count, n = x.shape
    rc = getattr(lib(), "or_fft_batch_" + _sfx(x.dtype))(n, count, _p(x), _p(out), int(inverse), int(nthreads),
                                                         int(exact_chirp))
    assert rc == 0
    return out


def fft_strided(x, n, stride, inverse=False):
    """Plan.ForwardStrided (plan_strided.go:35-99); x is the flat buffer."""
    x = np.ascontiguousarray(x)
    out = x.copy()
    rc = getattr(lib(), "or_fft_strided_" + _sfx(x.dtype))(n, stride, _p(x), _p(out), int(inverse))
    assert rc == 0
    return out


def rfft(x):
    """PlanRealT.Forward over rows (plan_real_generic.go:142-228)."""
    x = np.ascontiguousarray(x)
    cdt = np.complex64 if x.dtype == np.float32 else np.complex128
    x2 = x.reshape(-1, x.shape[-1])
    n = x2.shape[1]
    out = np.empty((x2.shape[0], n // 2 + 1), dtype=cdt)
    rc = getattr(lib(), "or_rfft_" + _sfx(cdt))(n, x2.shape[0], _p(x2), _p(out))
    assert rc == 0
    return out.reshape(x.shape[:-1] + (n // 2 + 1,))


def irfft(X, n):
    """PlanRealT.Inverse over rows (plan_real_generic.go:289-355); raises
    ValueError for the reference's ErrInvalidSpectrum."""
    X = np.ascontiguousarray(X)
    rdt = np.float32 if X.dtype == np.complex64 else np.float64
    X2 = X.reshape(-1, X.shape[-1])
    out = np.empty((X2.shape[0], n), dtype=rdt)
    rc = getattr(lib(), "or_irfft_" + _sfx(X.dtype))(n, X2.shape[0], _p(X2), _p(out))
    if rc == -2:
        raise ValueError("invalid spectrum")
    assert rc == 0
    return out.reshape(X.shape[:-1] + (n,))


def fftn(x, inverse=False):
    """PlanND (plan_nd.go:331-415)."""
    x = np.ascontiguousarray(x)
    dims = np.array(x.shape, dtype=np.int64)
    out = np.empty_like(x)
    rc = getattr(lib(), "or_fftnd_" + _sfx(x.dtype))(x.ndim, _p(dims), _p(x), _p(out), int(inverse))
    assert rc == 0
    return out


def rfftn(x):
    """PlanReal2D/PlanReal3D forward (plan_real_3d.go:145-226)."""
    x = np.ascontiguousarray(x)
    cdt = np.complex64 if x.dtype == np.float32 else np.complex128
    dims = np.array(x.shape, dtype=np.int64)
    out = np.empty(x.shape[:-1] + (x.shape[-1] // 2 + 1,), dtype=cdt)
    rc = getattr(lib(), "or_rfftnd_" + _sfx(cdt))(x.ndim, _p(dims), _p(x), _p(out))
    assert rc == 0
    return out


def convolve(a, b):
    """Convolve (convolve.go:9-59)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    out = np.empty(a.size + b.size - 1, dtype=a.dtype)
    rc = getattr(lib(), "or_convolve_" + _sfx(a.dtype))(a.size, _p(a), b.size, _p(b), _p(out))
    assert rc == 0
    return out


def convolve_real(a, b):
    """ConvolveReal (convolve_real.go:9-80): float32 or float64 inputs."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    out = np.empty(a.size + b.size - 1, dtype=a.dtype)
    sfx = "c64" if a.dtype == np.float32 else "c128"
    rc = getattr(lib(), "or_convolve_real_" + sfx)(a.size, _p(a), b.size, _p(b), _p(out))
    assert rc == 0
    return out


def twiddles(n, dtype=np.complex128):
    """ComputeTwiddleFactors (internal/math/twiddle.go:8-21)."""
    out = np.empty(n, dtype=dtype)
    getattr(lib(), "or_twiddles_" + _sfx(dtype))(n, _p(out))
    return out


def chirp(n, exact=True, dtype=np.complex128):
    """ComputeChirpSequence (internal/kernels/bluestein.go:20-32)."""
    out = np.empty(n, dtype=dtype)
    getattr(lib(), "or_chirp_" + _sfx(dtype))(n, int(exact), _p(out))
    return out


def route(n):
    """Which algorithm the reference routes n to: 0 trivial, 1 stockham,
    2 sixstep, 3 mixed radix, 4 bluestein."""
    return getattr(lib(), "or_route_c128")(n)


def dft_naive(x, inverse=False):
    """Brute-force DFT in long double (internal/reference/dft.go:16-100)."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    rc = lib().or_dft_naive(x.size, _p(x), _p(out), int(inverse))
    assert rc == 0
    return out


def rng_u01(seed, count, offset=0):
    out = np.empty(count, dtype=np.float64)
    lib().or_rng_u01(seed, offset, count, _p(out))
    return out


def set_fault(f):
    """Plant a known mistake in the oracle (0 = none); see afft_oracle.c."""
    lib().or_set_fault(int(f))


def mlp_loss_grad(params, X, T):
    """FP64 loss and gradient of the cost-model MLP (checker for the GPU
    training kernels); X and T already standardised."""
    P = np.ascontiguousarray(params, np.float64)
    X = np.ascontiguousarray(X, np.float64)
    T = np.ascontiguousarray(T, np.float64)
    G = np.empty_like(P)
    loss = ctypes.c_double()
    L = lib()
    L.or_mlp_loss_grad.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long,
                                   ctypes.c_void_p, ctypes.c_void_p]
    rc = L.or_mlp_loss_grad(_p(P), _p(X), _p(T), len(T), ctypes.addressof(loss), _p(G))
    assert rc == 0
    return loss.value, G
