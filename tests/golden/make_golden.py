"""Writes the committed golden vectors used to pin the oracle.

    python tests/golden/make_golden.py

The expected outputs come from numpy.fft (pocketfft, FP64), an independent
library implementation of the same DFT definition X[k] = Σ x[n] e^{-2πink/N}
the reference uses (plan.go:325).  Inputs come from the portable seeded
stream of algofft_hip.synth, whose first values are also stored so the C
and Python generators stay pinned to each other.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "algo-fft_amd"))
from algofft_hip import synth  # noqa: E402

C2C_SIZES = [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 15, 16, 17, 31, 64, 100, 128, 243, 256, 343, 1000, 1009, 1024, 2048,
             4096]
R2C_SIZES = [2, 8, 30, 64, 1000, 1024]
ND_SHAPES = [(4, 6), (3, 5, 8), (8, 8, 8)]


def main():
    out = {}
    for n in C2C_SIZES:
        x = synth.uniform_complex(1000 + n, (n,), np.complex128)
        out[f"c2c_x_{n}"] = x
        out[f"c2c_X_{n}"] = np.fft.fft(x)
    for n in R2C_SIZES:
        x = synth.uniform_real(2000 + n, (n,), np.float64)
        out[f"r2c_x_{n}"] = x
        out[f"r2c_X_{n}"] = np.fft.rfft(x)
    for shp in ND_SHAPES:
        key = "x".join(map(str, shp))
        x = synth.uniform_complex(3000 + len(shp), shp, np.complex128)
        out[f"nd_x_{key}"] = x
        out[f"nd_X_{key}"] = np.fft.fftn(x)
    out["rng_seed0_first16"] = synth.u01(0, 16)
    out["rng_seed1234_first16"] = synth.u01(1234, 16)
    path = os.path.join(HERE, "numpy_fft_vectors.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()
