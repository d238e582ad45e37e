"""real_time:: mirror (src/real_time/): the ETRS step and a frozen-potential
propagation loop.

Perturbations follow INQ's duck-typed interface (perturbations/none.hpp:20-55):
an object with has_uniform_vector_potential() and uniform_vector_potential(t)
returning the Cartesian A(t); self_consistency converts it to covariant
components (self_consistency.hpp:152-156).
"""
import ctypes

import numpy as np

from . import _lib
from .hamiltonian import _orbital_args


def _to_covariant(lattice, A_cart):
    # cell_metric::to_covariant(cartesian) (cell.hpp:184-187)
    return np.asarray(lattice, float) @ np.asarray(A_cart, float)


def etrs(ham, phi, time, dt, pert=None, order=4):
    """One ETRS step for a frozen potential (etrs.hpp:22-82):
    phi <- T(i dt/2, H(t+dt)) T(i dt/2, H(t)) phi.  Leaves H at A(t+dt)."""
    p, nst, ld = _orbital_args(phi)
    lat = ham.system.lattice
    if pert is not None and pert.has_uniform_vector_potential():
        A_t = _to_covariant(lat, pert.uniform_vector_potential(time))
        A_tdt = _to_covariant(lat, pert.uniform_vector_potential(time + dt))
    else:
        A_t = np.zeros(3)
        A_tdt = np.zeros(3)
    ham.handle.call("inqamd_etrs_step", ctypes.c_double(dt), _lib.dvec(A_t), _lib.dvec(A_tdt),
                    int(order), p, nst, ld)
    ham._A = np.asarray(A_tdt, float).copy()
    return phi


def propagate(ham, phi, dt, num_steps, pert=None, start_step=0, order=4, callback=None):
    """real_time::propagate's step loop (propagate.hpp:93-133) without the
    energy/observable bookkeeping; `callback(step, time, phi)` after each step."""
    for istep in range(start_step, num_steps):
        etrs(ham, phi, istep * dt, dt, pert, order)
        if callback is not None:
            callback(istep + 1, (istep + 1) * dt, phi)
    return phi


class fourier_series_vector_potential:
    """A_x(t) = sum_j c_j sin(j w t) (Cartesian, along `direction`)."""

    def __init__(self, coefficients, frequency, direction=(1.0, 0.0, 0.0)):
        self.c = np.asarray(coefficients, float)
        self.w = float(frequency)
        self.dir = np.asarray(direction, float)

    def has_uniform_vector_potential(self):
        return True

    def uniform_vector_potential(self, t):
        j = np.arange(1, self.c.size + 1)
        return self.dir * float(np.sum(self.c * np.sin(j * self.w * t)))
