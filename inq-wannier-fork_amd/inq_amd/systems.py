"""Synthetic Kohn-Sham systems of the frozen spec (spec/systems.json).

A system is the data ks_hamiltonian consumes: the grid, the lattice, the
local potential V(r) and the packed projector tables of projector_all
(src/hamiltonian/projector_all.hpp:408-419), plus seeded orthonormal
orbitals.  Everything here is host-side numpy; the geometry goes through the
library's C ABI builders so the FP32 sphere tables are the reference's.
"""
import json
import os
from dataclasses import dataclass, field

import numpy as np

from . import geometry

_SPEC_PATH = os.path.normpath(os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..",
                                           "spec", "systems.json"))
_spec_cache = None


def spec():
    global _spec_cache
    if _spec_cache is None:
        with open(_SPEC_PATH) as f:
            _spec_cache = json.load(f)
    return _spec_cache


def rng(seed, system_id, stream):
    """Counter-based generator keyed by (seed, system, stream)."""
    return np.random.Generator(np.random.Philox(key=[int(seed) & (2**64 - 1),
                                                     (int(system_id) << 8) | int(stream)]))


@dataclass
class System:
    n: tuple
    lattice: np.ndarray          # [3,3] rows a0, a1, a2 (bohr)
    vloc: np.ndarray             # [nx,ny,nz] (Ha)
    atoms: np.ndarray            # [natoms,3] cartesian (bohr)
    tables: dict = field(default_factory=dict)

    @property
    def npoints(self):
        return int(np.prod(self.n))

    @property
    def volume(self):
        return float(abs(np.linalg.det(self.lattice)))

    @property
    def natoms(self):
        return int(self.atoms.shape[0])


def symmetric_indices(n):
    i = np.arange(n)
    return np.where(i >= (n + 1) // 2, i - n, i)


def gaussian_vloc(lattice, n, atoms, depth, width):
    """V(r) = sum_atoms sum_images depth exp(-|r - R - L|^2 / (2 width^2)) on an
    orthorhombic cell, evaluated as a product of converged 1-D image sums.
    Grid point r = symmetric index * a/n (real_space.hpp:76-79)."""
    lat = np.asarray(lattice, float)
    if np.abs(lat - np.diag(np.diag(lat))).max() > 0:
        return _gaussian_vloc_general(lat, n, atoms, depth, width)
    L = np.diag(lat)
    coords = [symmetric_indices(n[d]) * (L[d] / n[d]) for d in range(3)]
    V = np.zeros(tuple(n))
    for R in np.asarray(atoms, float):
        fac = []
        for d in range(3):
            nimg = int(np.ceil(12.0 * width / L[d])) + 1
            x = coords[d][:, None] - R[d] - L[d] * np.arange(-nimg, nimg + 1)[None, :]
            fac.append(np.exp(-x * x / (2.0 * width * width)).sum(axis=1))
        V += depth * fac[0][:, None, None] * fac[1][None, :, None] * fac[2][None, None, :]
    return V


def _gaussian_vloc_general(lat, n, atoms, depth, width):
    """Direct image sum for a general (non-orthogonal) cell; for small grids."""
    c = [symmetric_indices(n[d]) / n[d] for d in range(3)]
    C0, C1, C2 = np.meshgrid(*c, indexing="ij")
    r = C0[..., None] * lat[0] + C1[..., None] * lat[1] + C2[..., None] * lat[2]
    reach = 12.0 * width
    nimg = [int(np.ceil(reach / np.linalg.norm(lat[d]))) + 1 for d in range(3)]
    V = np.zeros(tuple(n))
    for R in np.asarray(atoms, float):
        for i in range(-nimg[0], nimg[0] + 1):
            for j in range(-nimg[1], nimg[1] + 1):
                for k in range(-nimg[2], nimg[2] + 1):
                    d = r - (R + i * lat[0] + j * lat[1] + k * lat[2])
                    V += depth * np.exp(-(d * d).sum(-1) / (2.0 * width * width))
    return V


def projector_tables(lattice, n, atoms, radius, sigma, nlm, coeff):
    """Packed projector_all tables for Gaussian projectors on every atom."""
    spheres = [geometry.spherical_grid(lattice, n, R, radius) for R in np.asarray(atoms, float)]
    nproj = len(spheres)
    K = np.array([s[1].size for s in spheres], np.int32)
    maxK = int(K.max()) if nproj else 0
    points = -np.ones((nproj, maxK, 3), np.int32)
    positions = np.zeros((nproj, maxK, 3), np.float32)
    matrices = np.zeros((nproj, nlm, maxK), np.float64)
    for p, (c, d, r) in enumerate(spheres):
        k = d.size
        points[p, :k] = c
        positions[p, :k] = r
        matrices[p, :, :k] = geometry.gaussian_projector_matrix(lattice, d, r, sigma, nlm)
    return {
        "nproj": nproj, "max_sphere": maxK, "max_nlm": nlm,
        "points": points, "positions": positions, "matrices": matrices,
        "coeff": np.ascontiguousarray(coeff, np.float64).reshape(nproj, nlm),
        "nlm": np.full(nproj, nlm, np.int32), "sphere_size": K,
    }


def diamond_atoms(supercell, a0):
    basis = np.array(spec()["atoms"]["basis_fractional"], float)
    m = int(supercell)
    shifts = np.array([[i, j, k] for i in range(m) for j in range(m) for k in range(m)], float)
    frac = (basis[None, :, :] + shifts[:, None, :]).reshape(-1, 3) / m
    return frac * (a0 * m)


def make_system(n, lattice, atoms, seed=0, system_id=0, with_projectors=True):
    sp = spec()
    V = gaussian_vloc(lattice, n, atoms, sp["vloc"]["depth_ha"], sp["vloc"]["width_bohr"])
    sysm = System(n=tuple(int(x) for x in n), lattice=np.asarray(lattice, float), vloc=V,
                  atoms=np.asarray(atoms, float))
    if with_projectors and len(atoms) > 0:
        pj = sp["projectors"]
        nlm = int(pj["nlm"])
        lo, hi = pj["coeff_range_ha"]
        coeff = rng(seed, system_id, 1).uniform(lo, hi, size=(len(atoms), nlm))
        sysm.tables = projector_tables(lattice, n, atoms, pj["sphere_radius_bohr"],
                                       pj["sigma_bohr"], nlm, coeff)
    else:
        sysm.tables = {"nproj": 0}
    return sysm


def diamond_system(supercell=1, grid=None, seed=0, system_id=0, with_projectors=True):
    """Config systems: a = supercell * a0, grid = supercell * 32 unless given."""
    sp = spec()
    a0 = sp["cell"]["a0_bohr"]
    a = a0 * supercell
    g = int(grid) if grid is not None else sp["cell"]["points_per_a0"] * supercell
    lattice = np.eye(3) * a
    return make_system((g, g, g), lattice, diamond_atoms(supercell, a0), seed, system_id,
                       with_projectors)


def uneven_batch_layout(seed=None):
    """(grid, bands, atoms) of every system of the uneven batch, all drawn from
    one generator keyed (seed, 1000, 3): grid uniform over the listed sizes,
    bands round(exp(U(ln lo, ln hi))), atoms uniform integers in [lo, hi]."""
    sp = spec()
    cfg = sp["configs"]["uneven_batch"]
    seed = sp["seed"] if seed is None else seed
    g = rng(seed, 1000, 3)
    grids = np.array(cfg["grids"])
    blo, bhi = cfg["bands_log_uniform"]
    alo, ahi = cfg["atoms_uniform"]
    out = []
    for _ in range(cfg["systems"]):
        L = int(g.choice(grids))
        nst = int(round(np.exp(g.uniform(np.log(blo), np.log(bhi)))))
        na = int(g.integers(alo, ahi + 1))
        out.append((L, nst, na))
    return out


def uneven_batch_system(index, grid, natoms, seed=None, with_projectors=True):
    """System `index` of the uneven batch: a cubic cell a = grid * h (the
    config cells' spacing h = a0 / points_per_a0), atoms at uniform fractional
    positions from (seed, 100 + index, 2)."""
    sp = spec()
    seed = sp["seed"] if seed is None else seed
    h = sp["cell"]["a0_bohr"] / sp["cell"]["points_per_a0"]
    lattice = np.eye(3) * (grid * h)
    frac = rng(seed, 100 + index, 2).uniform(0.0, 1.0, size=(natoms, 3))
    return make_system((grid, grid, grid), lattice, frac @ lattice, seed, 100 + index,
                       with_projectors)


def plane_wave_mask(sysm, ecut):
    """Boolean [nx,ny,nz] of G with |G|^2/2 <= ecut (G = sum_d m_d b_d)."""
    b = 2.0 * np.pi * np.linalg.inv(sysm.lattice).T
    m = [symmetric_indices(sysm.n[d]) for d in range(3)]
    M0, M1, M2 = np.meshgrid(*m, indexing="ij")
    G = M0[..., None] * b[0] + M1[..., None] * b[1] + M2[..., None] * b[2]
    return 0.5 * (G * G).sum(axis=-1) <= ecut


def random_orbitals(sysm, nst, seed=0, system_id=0, ecut=None, fourier=False):
    """Seeded orthonormal orbitals; returns psi(r) [nx,ny,nz,nst] complex128
    (or psi(G), unnormalised-FFT convention, when fourier=True)."""
    if ecut is None:
        ecut = spec()["orbitals"]["ecut_ha"]
    mask = plane_wave_mask(sysm, ecut)
    npw = int(mask.sum())
    if npw < nst:
        raise ValueError(f"only {npw} plane waves for {nst} bands")
    g = rng(seed, system_id, 0)
    C = g.standard_normal((nst, npw)) + 1j * g.standard_normal((nst, npw))
    N = sysm.npoints
    dvg = sysm.volume / (N * N)  # fourier_space volume element (fourier_space.hpp:41-44)
    # orthogonalize (orthogonalize.hpp:25-32): Cholesky of the overlap
    # S = L L^H with S[a,b] = dvg sum_G C[a,G] conj(C[b,G]); C <- L^-1 C gives S = 1
    S = dvg * (C @ C.conj().T)
    Lc = np.linalg.cholesky(S)
    C = np.linalg.solve(Lc, C)
    psi_g = np.zeros(sysm.n + (nst,), np.complex128)
    psi_g[mask] = C.T
    if fourier:
        return psi_g
    try:
        import scipy.fft as sfft
        return sfft.ifftn(psi_g, axes=(0, 1, 2), workers=-1)
    except ImportError:
        return np.fft.ifftn(psi_g, axes=(0, 1, 2))
