"""Pins the oracle's FFT (operations::transform restatement).

Checks: numpy.fft golden vectors (committed fixtures), an O(N^2) brute-force
DFT on tiny grids, the reference's Gaussian transform and round-trip tests
(src/operations/transform.hpp:442-545), Parseval with INQ's volume elements
(real_space.hpp:150-152, fourier_space.hpp:41-44), and a plane wave mapping
to a single G.  A wrong sign, a dropped twiddle or a transposed axis fails
the golden and brute-force checks.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from helpers import grid_positions, reciprocal, relerr, symmetric

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "fft_vectors.npz"))
KNOWN = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))


@pytest.mark.parametrize("key", sorted(k[len("fft3_in_"):] for k in GOLD.files if k.startswith("fft3_in_")))
def test_fft3_golden(key):
    a = GOLD[f"fft3_in_{key}"]
    assert relerr(O.to_fourier(a), GOLD[f"fft3_fwd_{key}"]) < 1e-14
    assert relerr(O.to_real(a, normalize=False), GOLD[f"fft3_bwd_{key}"]) < 1e-14


@pytest.mark.parametrize("n", range(1, 41))
def test_fft1_golden(n):
    v = GOLD[f"fft1_in_{n}"]
    assert relerr(O.fft1d(v, -1), GOLD[f"fft1_fwd_{n}"]) < 1e-14
    assert relerr(O.fft1d(v, +1), GOLD[f"fft1_bwd_{n}"]) < 1e-14


@pytest.mark.parametrize("g", [(2, 3, 5), (6, 1, 4), (7, 3, 2)])
def test_fft3_brute_force(g):
    rng = np.random.default_rng(1)
    a = rng.standard_normal(g + (2,)) + 1j * rng.standard_normal(g + (2,))
    idx = [np.arange(n) for n in g]
    ref = np.zeros_like(a)
    for kx in range(g[0]):
        for ky in range(g[1]):
            for kz in range(g[2]):
                ph = np.exp(-2j * np.pi * (np.multiply.outer(np.multiply.outer(kx * idx[0] / g[0], np.ones(g[1])), np.ones(g[2]))
                                           + np.multiply.outer(np.multiply.outer(np.ones(g[0]), ky * idx[1] / g[1]), np.ones(g[2]))
                                           + np.multiply.outer(np.ones((g[0], g[1])), kz * idx[2] / g[2])))
                ref[kx, ky, kz] = np.einsum("xyz,xyzb->b", ph, a)
    assert relerr(O.to_fourier(a), ref) < 1e-13


def test_transform_zero_and_gaussian_round_trip():
    """transform.hpp:442-545: zero maps to zero; Gaussian round trip < 1e-15."""
    t = KNOWN["transform"]
    L = t["cell"]
    n = int(round(L / t["spacing"]))
    lat = np.eye(3) * L
    r = grid_positions(lat, (n, n, n))
    r2 = (r * r).sum(-1)
    nst = t["bands"]
    zero = np.zeros((n, n, n, nst), complex)
    assert np.abs(O.to_fourier(zero)).mean() < 1e-15
    phi = np.stack([np.exp(-0.5 * (b + 1) * r2) for b in range(nst)], axis=-1).astype(complex)
    back = O.to_real(O.to_fourier(phi))
    assert np.abs(back - phi).mean() < t["round_trip_mean_abs"]


def test_parseval_volume_elements():
    rng = np.random.default_rng(3)
    n = (6, 8, 10)
    lat = np.diag([3.0, 4.0, 5.0])
    vol = 60.0
    N = np.prod(n)
    a = rng.standard_normal(n + (2,)) + 1j * rng.standard_normal(n + (2,))
    f = O.to_fourier(a)
    lhs = (vol / N) * (np.abs(a) ** 2).sum(axis=(0, 1, 2))
    rhs = (vol / N ** 2) * (np.abs(f) ** 2).sum(axis=(0, 1, 2))
    assert np.allclose(lhs, rhs, rtol=1e-13)


def test_plane_wave_single_g():
    n = (8, 6, 10)
    lat = np.diag([4.0, 3.0, 5.0])
    b = reciprocal(lat)
    m = (3, -2, 4)
    G = m[0] * b[0] + m[1] * b[1] + m[2] * b[2]
    r = grid_positions(lat, n)
    phi = np.exp(1j * (r @ G))[..., None]
    f = O.to_fourier(phi)[..., 0]
    idx = tuple(mm % nn for mm, nn in zip(m, n))
    assert abs(f[idx] - np.prod(n)) < 1e-10
    f[idx] = 0
    assert np.abs(f).max() < 1e-10
