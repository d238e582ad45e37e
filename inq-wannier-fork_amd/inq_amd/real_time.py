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
from .hamiltonian import _orbital_args, _stream_ptr


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


class etrs_graph:
    """One ETRS step on a fixed orbital buffer captured as a hipGraph.

    With `pert=None` each `launch(time)` needs the covariant A(t), A(t+dt)
    from the host (written to device memory before the replay); with a
    fourier_series_vector_potential the graph evaluates A itself from a device
    step counter and `launch()` takes no arguments.
    """

    def __init__(self, ham, phi, dt, order=4, pert=None):
        p, nst, ld = _orbital_args(phi)
        self._ham = ham
        self._phi = phi
        self.dt = float(dt)
        self.pert = pert
        ser = None
        if isinstance(pert, fourier_series_vector_potential):
            self._c = np.ascontiguousarray(pert.c, np.float64)
            ser = _lib.VectorPotentialSeries()
            ser.n = int(self._c.size)
            ser.coeff = self._c.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            ser.omega = float(pert.w)
            ser.direction = (ctypes.c_double * 3)(*[float(x) for x in pert.dir])
        g = ctypes.c_void_p()
        ham.handle.call("inqamd_etrs_graph_create", p, nst, ld, ctypes.c_double(dt), int(order),
                        ctypes.byref(ser) if ser is not None else None, ctypes.byref(g))
        self._g = g
        self.series = ser is not None

    def launch(self, time=None):
        h = self._ham.handle
        _lib.check(h._lib.inqamd_set_stream(h.h, _stream_ptr()), h.h)
        if self.series:
            rc = h._lib.inqamd_etrs_graph_launch(self._g, None, None)
        else:
            lat = self._ham.system.lattice
            if self.pert is not None and self.pert.has_uniform_vector_potential():
                A_t = _to_covariant(lat, self.pert.uniform_vector_potential(time))
                A_tdt = _to_covariant(lat, self.pert.uniform_vector_potential(time + self.dt))
            else:
                A_t = A_tdt = np.zeros(3)
            rc = h._lib.inqamd_etrs_graph_launch(self._g, _lib.dvec(A_t), _lib.dvec(A_tdt))
            self._ham._A = np.asarray(A_tdt, float).copy()
        _lib.check(rc, h.h)

    def set_step(self, step):
        _lib.check(self._ham.handle._lib.inqamd_etrs_graph_set_step(self._g, int(step)),
                   self._ham.handle.h)

    def close(self):
        if getattr(self, "_g", None):
            _lib.lib().inqamd_graph_destroy(self._g)
            self._g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


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
