"""Ground-state eigensolver pieces on G-space orbitals.

Mirrors operations::laplacian_add (src/operations/laplacian.hpp:24-38),
operations::preconditioner (src/operations/preconditioner.hpp:25-55),
operations::orthogonalize (src/operations/orthogonalize.hpp:25-32),
eigensolvers::steepest_descent (src/eigensolvers/steepest_descent.hpp:23-88)
and ground_state::subspace_diagonalization
(src/ground_state/subspace_diagonalization.hpp:16-25), as
ground_state::calculator uses them (calculator.hpp:150-180): a subspace
diagonalisation, the orbitals to G space, one steepest-descent call per SCF
iteration, then back.  All run in libinq_amd.so on the GPU.
"""
import ctypes

import numpy as np

from . import _lib
from .hamiltonian import _orbital_args


def laplacian_add(ham, ff, laplff, factor=1.0, gradcoeff=(0.0, 0.0, 0.0)):
    """laplff += factor (-|G|^2 + gradcoeff . G) ff on G-space arrays, in place
    (gradcoeff contravariant, G covariant = 2 pi m)."""
    f, nst, ld = _orbital_args(ff)
    o, nst2, _ = _orbital_args(laplff)
    if nst2 != nst or ff.shape != laplff.shape:
        raise ValueError("ff and laplff must have the same shape")
    ham.handle.call("inqamd_laplacian_add", f, o, nst, ld, ctypes.c_double(factor),
                    _lib.dvec(gradcoeff))
    return laplff


def preconditioner(ham, phi):
    """operations::preconditioner on G-space orbitals, in place."""
    p, nst, ld = _orbital_args(phi)
    ham.handle.call("inqamd_preconditioner", p, nst, ld)
    return phi


def orthogonalize(ham, phi, fourier=False):
    """operations::orthogonalize (Gram-Schmidt in band order through a
    Cholesky factor of the overlap), in place; `fourier` selects the G-space
    volume element V/N^2 instead of V/N."""
    p, nst, ld = _orbital_args(phi)
    ham.handle.call("inqamd_orthogonalize", p, nst, ld, int(bool(fourier)))
    return phi


def steepest_descent(ham, phi, precondition=True):
    """eigensolvers::steepest_descent on G-space orbitals with the Hamiltonian's
    G-space operator (and the TPA preconditioner), in place."""
    p, nst, ld = _orbital_args(phi)
    ham.handle.call("inqamd_steepest_descent", p, nst, ld, int(bool(precondition)))
    return phi


def subspace_diagonalization(ham, phi):
    """ground_state::subspace_diagonalization on real-space orbitals, in place;
    returns the subspace eigenvalues (ascending, numpy array)."""
    p, nst, ld = _orbital_args(phi)
    eig = np.zeros(nst)
    ham.handle.call("inqamd_subspace_diagonalization", p, nst, ld,
                    eig.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return eig


def eigensolve(ham, phi, iterations, precondition=True, subspace_diag=True):
    """A frozen-potential ground-state loop (calculator.hpp:150-180 without the
    density update): per iteration the subspace diagonalisation (INQ's
    default), to_fourier, one steepest-descent call, to_real.  Returns the new
    real-space orbitals and, per iteration, the energy::calculate eigenvalues
    and residual norms (device tensors)."""
    from .operations import transform
    phi = phi.clone()
    history = []
    for _ in range(iterations):
        if subspace_diag:
            subspace_diagonalization(ham, phi)
        fphi = transform.to_fourier(phi, ham)
        steepest_descent(ham, fphi, precondition)
        transform.to_real(fphi, ham, out=phi)
        eps, normres, _ = ham.energy(phi)
        history.append((eps.clone(), normres.clone()))
    return phi, history
