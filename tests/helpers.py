"""Shared helpers for the tests (no GPU access here)."""
import numpy as np


def rel_l2(got, ref):
    got = np.asarray(got, dtype=np.complex128)
    ref = np.asarray(ref, dtype=np.complex128)
    den = np.linalg.norm(ref.ravel())
    return float(np.linalg.norm((got - ref).ravel()) / (den if den > 0 else 1.0))


def max_rel(got, ref):
    got = np.asarray(got, dtype=np.complex128)
    ref = np.asarray(ref, dtype=np.complex128)
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))


def c64_tol(n):
    """complex64 budget against the FP64 oracle: rel L2 <= 2e-6 * log2(N)."""
    return 2e-6 * max(1.0, float(np.log2(max(n, 2))))


C128_TOL = 1e-13
