"""Shared test helpers (error measures, small systems)."""
import numpy as np


def relerr(a, b):
    """||a - b||_2 / ||b||_2 over the whole array."""
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def band_relerr(a, b):
    """max over bands (last axis) of ||a_b - b_b|| / ||b_b||."""
    a = np.asarray(a).reshape(-1, np.shape(a)[-1])
    b = np.asarray(b).reshape(-1, np.shape(b)[-1])
    num = np.linalg.norm(a - b, axis=0)
    den = np.maximum(np.linalg.norm(b, axis=0), 1e-300)
    return float((num / den).max())


def symmetric(n):
    i = np.arange(n)
    return np.where(i >= (n + 1) // 2, i - n, i)


def grid_positions(lattice, n):
    """Cartesian r of every grid point, [nx,ny,nz,3] (real_space.hpp:76-95)."""
    lat = np.asarray(lattice, float)
    c = [symmetric(n[d]) / n[d] for d in range(3)]
    C0, C1, C2 = np.meshgrid(*c, indexing="ij")
    return C0[..., None] * lat[0] + C1[..., None] * lat[1] + C2[..., None] * lat[2]


def reciprocal(lattice):
    return 2.0 * np.pi * np.linalg.inv(np.asarray(lattice, float)).T
