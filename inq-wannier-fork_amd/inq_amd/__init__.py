"""inq_amd -- MI355X (gfx950) plane-wave Hpsi and real-time propagator.

Python mirror of INQ's hamiltonian / operations / real_time interfaces over
the C ABI of libinq_amd.so (include/inq_amd.h).  The host-side geometry and
synthetic-system builders need no GPU; everything that computes on orbitals
runs the HIP library and fails loudly if it is absent.
"""
from . import _lib, geometry, systems  # noqa: F401
from .hamiltonian import ks_hamiltonian, projector_all  # noqa: F401
from . import operations, real_time, observables, adjoint, batch, ground_state, tddft, checkpoint  # noqa: F401
from .operations import transform, space, fourier_plan  # noqa: F401

__all__ = ["ks_hamiltonian", "projector_all", "operations", "real_time", "transform", "space",
           "fourier_plan", "geometry", "systems", "observables", "adjoint", "batch", "ground_state", "tddft", "checkpoint"]
