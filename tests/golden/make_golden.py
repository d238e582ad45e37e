"""Writes the golden vectors under tests/golden/ (run from the repo root).

fft_vectors.npz: inputs and numpy.fft outputs (an independent library, not
INQ code) for batched 3-D transforms over the first three axes and for 1-D
transforms of every length 1..40 (radix 2/3/5 and primes up to 37).
Seeds are fixed; rerunning reproduces the file bit for bit.
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20241202)
    out = {}
    grids = [(6, 6, 6), (10, 12, 15), (17, 5, 7), (8, 9, 10), (12, 12, 12), (4, 30, 3)]
    for g in grids:
        nst = 3
        a = rng.standard_normal(g + (nst,)) + 1j * rng.standard_normal(g + (nst,))
        key = "x".join(map(str, g))
        out[f"fft3_in_{key}"] = a
        out[f"fft3_fwd_{key}"] = np.fft.fftn(a, axes=(0, 1, 2))
        out[f"fft3_bwd_{key}"] = np.fft.ifftn(a, axes=(0, 1, 2)) * np.prod(g)
    for n in range(1, 41):
        v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"fft1_in_{n}"] = v
        out[f"fft1_fwd_{n}"] = np.fft.fft(v)
        out[f"fft1_bwd_{n}"] = np.fft.ifft(v) * n
    np.savez_compressed(os.path.join(HERE, "fft_vectors.npz"), **out)


if __name__ == "__main__":
    main()
