"""Writes tests/golden/fft_vectors.npz: seeded inputs and numpy (pocketfft)
FP64 outputs for the 1-D transform contract of epfftw.f (FFTW sign
convention, unnormalised): c2c forward/backward, r2c, c2r.

    python tests/golden/make_fft_vectors.py
"""
import pathlib

import numpy as np

out = {}
rng = np.random.default_rng(20240601)
for n in (8, 16, 32, 64):
    x = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    out[f"c2c_in_{n}"] = x
    out[f"c2c_fwd_{n}"] = np.fft.fft(x, axis=-1)            # exp(-i ...)
    out[f"c2c_bwd_{n}"] = np.fft.ifft(x, axis=-1) * n       # exp(+i ...), unnormalised
    r = rng.standard_normal((3, n))
    out[f"r2c_in_{n}"] = r
    out[f"r2c_out_{n}"] = np.fft.rfft(r, axis=-1)
    h = np.fft.rfft(rng.standard_normal((3, n)), axis=-1)
    out[f"c2r_in_{n}"] = h
    out[f"c2r_out_{n}"] = np.fft.irfft(h, n=n, axis=-1) * n
np.savez_compressed(pathlib.Path(__file__).with_name("fft_vectors.npz"), **out)
print("wrote", len(out), "arrays")
