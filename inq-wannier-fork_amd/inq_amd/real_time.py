"""real_time:: mirror (src/real_time/): the ETRS step, the propagation loop
with its per-step energy and observables, and hamiltonian::energy.

Perturbations follow INQ's duck-typed interface (perturbations/none.hpp:20-55):
an object with has_uniform_vector_potential() and uniform_vector_potential(t)
returning the Cartesian A(t); self_consistency converts it to covariant
components (self_consistency.hpp:152-156).  An optional zero_step(phi) runs
before a propagation that starts at step 0 (propagate.hpp:63-65).
"""
import ctypes

import numpy as np
import torch

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


class energy:
    """hamiltonian::energy (energy.hpp:19-200): the terms the propagation
    tracks; kinetic() and total() combine them as the reference does."""

    def __init__(self):
        self.ion = 0.0
        self.ion_kinetic = 0.0
        self.eigenvalues = 0.0
        self.external = 0.0
        self.non_local = 0.0
        self.hartree = 0.0
        self.xc = 0.0
        self.nvxc = 0.0
        self.exact_exchange = 0.0

    def calculate(self, ham, phi, occ):
        """energy::calculate (energy.hpp:59-107): one H application for the
        eigenvalues and residuals, the non-local expectation values, and the
        occupation-weighted sums.  Returns the per-band residual norms."""
        nst = phi.shape[-1]
        w = torch.as_tensor(np.broadcast_to(np.asarray(occ, np.float64), (nst,)).copy(),
                            dtype=torch.float64, device=phi.device)
        eps, normres, nl = ham.energy(phi)
        self.eigenvalues = float((w * eps).sum())
        self.non_local = float((w * nl).sum())
        self.exact_exchange = 0.0
        self.band_eigenvalues = eps
        return normres

    def kinetic(self):
        return (self.eigenvalues - 2.0 * self.hartree - self.nvxc - 2.0 * self.exact_exchange
                - self.external - self.non_local)

    def total(self):
        return (self.kinetic() + self.hartree + self.external + self.non_local + self.xc
                + self.exact_exchange + self.ion + self.ion_kinetic)


class viewables:
    """What a propagation callback sees (viewables.hpp:27-130)."""

    def __init__(self, last_iter, it, time, ham, phi, occ, en, pert):
        self._last = bool(last_iter)
        self._iter = int(it)
        self._time = float(time)
        self._ham = ham
        self._phi = phi
        self._occ = occ
        self._energy = en
        self._pert = pert

    def iter(self):
        return self._iter

    def last_iter(self):
        return self._last

    def every(self, every_iter):
        if self._iter == 0:
            return False
        return self._iter % every_iter == 0 or self._last

    def time(self):
        return self._time

    def energy(self):
        return self._energy

    def orbitals(self):
        return self._phi

    def uniform_vector_potential(self):
        """Cartesian A of the Hamiltonian (viewables.hpp:94-96)."""
        lat = np.asarray(self._ham.system.lattice, float)
        return np.linalg.solve(lat, self._ham.uniform_vector_potential)

    def current(self):
        """Cartesian total current (viewables.hpp:100-102) with the
        Hamiltonian's position_commutator form; bands of equal occupation."""
        occ = np.asarray(self._occ, float)
        if occ.ndim and not np.all(occ == occ.flat[0]):
            raise ValueError("current() needs equally occupied bands")
        return self._ham.current(self._phi, float(occ.flat[0]) if occ.ndim else float(occ))[1]

    def density(self):
        from . import tddft
        return tddft.density(self._ham, self._phi, self._occ)

    def dipole(self):
        """Electronic dipole of the density (observables::dipole, dipole.hpp:24-40)."""
        from . import tddft
        return tddft.dipole(self._ham, self.density())

    def num_electrons(self):
        rho = self.density()
        return float(rho.sum()) * self._ham.system.volume / rho.numel()


def _set_vector_potential(ham, pert, time):
    # self_consistency::update_hamiltonian's vector-potential part (:152-156)
    if pert is not None and pert.has_uniform_vector_potential():
        ham.uniform_vector_potential = _to_covariant(ham.system.lattice,
                                                     pert.uniform_vector_potential(time))
    else:
        ham.uniform_vector_potential = np.zeros(3)


def propagate(ham, phi, dt, num_steps, pert=None, start_step=0, order=4, callback=None, occ=2.0,
              theory=None, vion=None, ion_energy=0.0):
    """real_time::propagate (propagate.hpp:33-133) with ETRS for static ions.

    Before the loop the Hamiltonian is brought to t = start_step dt (A(t)
    and, with a theory, V_KS of phi's density) and the energy is calculated;
    a propagation from step 0 runs pert.zero_step(phi) first and reports
    step 0 to the callback.  Each step then runs etrs (the self-consistent
    one when `theory` is given), energy.calculate (one more H application)
    and callback(viewables).  `theory` is None (frozen potential, the
    Hamiltonian's V as the ionic one) or a dict(hartree=bool, xc=False |
    'lda' | 'pbe') with `vion` the ionic potential (a float64 device field).
    `occ` is one occupation or one per band.  Returns the final energy.
    """
    from . import tddft
    if start_step < 0:
        raise ValueError("start_step must be >= 0")
    nst = phi.shape[-1]
    occ_arr = np.broadcast_to(np.asarray(occ, np.float64), (nst,)).copy()
    dV = ham.system.volume / float(np.prod(ham.n))
    en = energy()
    en.ion = float(ion_energy)
    interacting = theory is not None
    if interacting:
        if vion is None:
            raise ValueError("an interacting propagation needs the ionic potential vion")
        hartree = bool(theory.get("hartree", True))
        xc = theory.get("xc", "pbe")
    if start_step == 0 and pert is not None and hasattr(pert, "zero_step"):
        pert.zero_step(phi)

    def update_hamiltonian(time):
        _set_vector_potential(ham, pert, time)
        rho = tddft.density(ham, phi, occ_arr)
        if interacting:
            e = tddft.ks_potential(ham, vion, rho, hartree, xc)
            en.external, en.hartree, en.xc, en.nvxc = (e["external"], e["hartree"], e["exc"],
                                                       e["nvxc"])
        else:
            v = ham.scalar_potential(phi.device)
            en.external = dV * float((rho.reshape(-1) * v).sum())

    update_hamiltonian(start_step * dt)
    en.calculate(ham, phi, occ_arr)
    if start_step == 0 and callback is not None:
        callback(viewables(False, start_step, start_step * dt, ham, phi, occ_arr, en, pert))
    for istep in range(start_step, num_steps):
        if interacting:
            tddft.etrs(ham, phi, vion, istep * dt, dt, occ_arr, pert, order, hartree, xc)
        else:
            etrs(ham, phi, istep * dt, dt, pert, order)
        # the step leaves H at t + dt; its energy terms are those of the new density
        update_hamiltonian((istep + 1) * dt)
        en.calculate(ham, phi, occ_arr)
        if callback is not None:
            callback(viewables(istep == num_steps - 1, istep + 1, (istep + 1.0) * dt, ham, phi,
                               occ_arr, en, pert))
    return en


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
