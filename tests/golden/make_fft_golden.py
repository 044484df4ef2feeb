"""Writes fft_golden.npz: seeded complex inputs and their DFTs from numpy.fft
(pocketfft), used to pin the oracle FFT and, through it, the device FFT.
Run from the repo root: python tests/golden/make_fft_golden.py"""
import os

import numpy as np

rng = np.random.default_rng(20261202)
out = {}
for n in (8, 12, 16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256, 7, 11, 25):
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    out[f"x1_{n}"] = x
    out[f"fwd1_{n}"] = np.fft.fft(x)
    out[f"inv1_{n}"] = np.fft.ifft(x) * n  # unnormalised inverse
for shp in ((4, 6, 8), (8, 12, 16), (16, 8, 24)):
    x = rng.normal(size=shp) + 1j * rng.normal(size=shp)   # (nz, ny, nx), x fastest
    k = "x".join(map(str, shp))
    out[f"x3_{k}"] = x
    out[f"fwd3_{k}"] = np.fft.fftn(x)
    out[f"inv3_{k}"] = np.fft.ifftn(x) * x.size
np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_golden.npz"), **out)
print("wrote", len(out), "arrays")
