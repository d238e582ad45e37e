"""operations:: mirror (src/operations/): transform (alias space), shift,
exponential_in_place, exponential, exponential_2_for_1.

`plan` is any object carrying a library handle for the grid: a
ks_hamiltonian or a fourier_plan.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .hamiltonian import _Handle, _orbital_args


class fourier_plan:
    """Grid-only handle (no potential, no projectors) for transforms."""

    def __init__(self, n, lattice, max_bands=1):
        self._h = _Handle(n, lattice, np.zeros(int(np.prod(n))), None, max_bands)
        self.n = tuple(int(x) for x in n)

    @property
    def handle(self):
        return self._h


class transform:
    """operations::transform (src/operations/transform.hpp:337-400)."""

    @staticmethod
    def to_fourier(phi, plan, out=None):
        p, nst, ld = _orbital_args(phi)
        if out is None:
            out = torch.empty_like(phi)
        q, _, _ = _orbital_args(out)
        plan.handle.call("inqamd_transform_to_fourier", p, q, nst, ld)
        return out

    @staticmethod
    def to_real(fphi, plan, normalize=True, out=None):
        p, nst, ld = _orbital_args(fphi)
        if out is None:
            out = torch.empty_like(fphi)
        q, _, _ = _orbital_args(out)
        plan.handle.call("inqamd_transform_to_real", p, q, nst, ld, int(bool(normalize)))
        return out


# operations::space: the same transforms under their other name
space = transform


def shift(factor, x, phi, plan):
    """phi += factor * x (src/operations/shift.hpp:33-44)."""
    p, nst, ld = _orbital_args(phi)
    q, _, _ = _orbital_args(x)
    c = complex(factor)
    plan.handle.call("inqamd_shift", ctypes.c_double(c.real), ctypes.c_double(c.imag), q, p, nst, ld)
    return phi


def exponential_in_place(ham, factor, phi, order=4):
    """src/operations/exponential.hpp:21-34."""
    p, nst, ld = _orbital_args(phi)
    z = complex(factor)
    ham.handle.call("inqamd_exponential_in_place", ctypes.c_double(z.real), ctypes.c_double(z.imag),
                    int(order), p, nst, ld)
    return phi


def exponential(ham, factor, phi, order=4):
    """src/operations/exponential.hpp:38-46."""
    expphi = phi.clone()
    return exponential_in_place(ham, factor, expphi, order)


def exponential_2_for_1(ham, factor1, factor2, phi, order=4):
    """src/operations/exponential.hpp:50-69: phi <- T(factor2) phi; returns T(factor1) phi."""
    p, nst, ld = _orbital_args(phi)
    expphi = torch.empty_like(phi)
    z1, z2 = complex(factor1), complex(factor2)
    ham.handle.call("inqamd_exponential_2_for_1", ctypes.c_double(z1.real), ctypes.c_double(z1.imag),
                    ctypes.c_double(z2.real), ctypes.c_double(z2.imag), int(order), p,
                    ctypes.c_void_p(expphi.data_ptr()), nst, ld)
    return expphi


def overlap_diagonal_normalized(hphi, phi, ham):
    """Re <hphi_b|phi_b> / <phi_b|phi_b> (overlap_diagonal.hpp:144-185, real_part)."""
    return ham.band_energies(hphi, phi)
