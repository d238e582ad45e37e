"""INQ-shaped host API over the C ABI, on torch ROCm tensors.

Orbitals are complex128 device tensors laid out like INQ's orbital_set:
shape [nx, ny, nz, nst] (or [N, nst]), band index contiguous.

    ham = ks_hamiltonian(system, max_bands)      # ks_hamiltonian.hpp:79-89
    hphi = ham(phi)                               # ks_hamiltonian.hpp:132-152
    ham.uniform_vector_potential = A_cov          # ks_hamiltonian.hpp:216-221
    vnl = ham.non_local(phi)                      # ks_hamiltonian.hpp:104-128
    sphere = ham.projectors_all().project(phi)    # projector_all.hpp:98-216
    ham.projectors_all().apply(sphere, vnl)       # projector_all.hpp:220-232
    J_cov, J_cart = ham.current(phi, occ)         # observables/current.hpp:25-53
"""
import ctypes

import numpy as np
import torch

from . import _lib


def _stream_ptr():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _orbital_args(phi):
    if not isinstance(phi, torch.Tensor):
        raise TypeError("orbitals must be torch tensors")
    if phi.dtype != torch.complex128:
        raise TypeError("orbitals must be complex128")
    if not phi.is_cuda:
        raise ValueError("orbitals must live on the GPU")
    if not phi.is_contiguous():
        raise ValueError("orbitals must be contiguous ([points][bands], band fastest)")
    nst = phi.shape[-1]
    return ctypes.c_void_p(phi.data_ptr()), nst, nst


class _Handle:
    """Owns one inqamd_hamiltonian_t."""

    def __init__(self, n, lattice, vloc, tables, max_bands):
        L = _lib.lib()
        self._lib = L
        d = _lib.SystemDesc()
        d.n = (ctypes.c_int * 3)(*[int(x) for x in n])
        d.lattice = (ctypes.c_double * 9)(*np.ravel(np.asarray(lattice, float)))
        v = np.ascontiguousarray(np.ravel(vloc), np.float64)
        keep = [v]
        d.vloc = v.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        d.max_bands = int(max_bands)
        if tables and tables.get("nproj", 0) > 0:
            arrs = {
                "points": (np.int32, ctypes.c_int), "positions": (np.float32, ctypes.c_float),
                "matrices": (np.float64, ctypes.c_double), "coeff": (np.float64, ctypes.c_double),
                "nlm": (np.int32, ctypes.c_int), "sphere_size": (np.int32, ctypes.c_int),
            }
            for name, (dt, ct) in arrs.items():
                a = np.ascontiguousarray(tables[name], dt)
                keep.append(a)
                setattr(d, name, a.ctypes.data_as(ctypes.POINTER(ct)))
            d.nproj = int(tables["nproj"])
            d.max_sphere = int(tables["max_sphere"])
            d.max_nlm = int(tables["max_nlm"])
        h = ctypes.c_void_p()
        rc = L.inqamd_hamiltonian_create(ctypes.byref(d), _stream_ptr(), ctypes.byref(h))
        _lib.check(rc, None)
        self.h = h
        self.n = tuple(int(x) for x in n)
        self.max_bands = int(max_bands)
        del keep

    def call(self, name, *args):
        L = self._lib
        _lib.check(L.inqamd_set_stream(self.h, _stream_ptr()), self.h)
        _lib.check(getattr(L, name)(self.h, *args), self.h)

    def close(self):
        if getattr(self, "h", None):
            self._lib.inqamd_hamiltonian_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class projector_all:
    """projector_all view of a ks_hamiltonian (projector_all.hpp:23-421)."""

    def __init__(self, ham):
        self._ham = ham
        t = ham.tables
        self.nprojs = int(t.get("nproj", 0)) if t else 0
        self.max_sphere_size = int(t.get("max_sphere", 0)) if t else 0
        self.max_nlm = int(t.get("max_nlm", 0)) if t else 0

    def project(self, phi):
        p, nst, ld = _orbital_args(phi)
        out = torch.zeros((self.nprojs, self.max_sphere_size, nst), dtype=torch.complex128,
                          device=phi.device)
        if self.nprojs:
            self._ham._h.call("inqamd_projector_project", p, ctypes.c_void_p(out.data_ptr()), nst, ld)
        return out

    def apply(self, sphere, vnlphi):
        p, nst, ld = _orbital_args(vnlphi)
        if self.nprojs:
            s = sphere.contiguous()
            self._ham._h.call("inqamd_projector_apply", ctypes.c_void_p(s.data_ptr()), p, nst, ld)
        return vnlphi


class ks_hamiltonian:
    """Frozen-potential Kohn-Sham Hamiltonian on one GPU."""

    def __init__(self, system, max_bands, kpoint=(0.0, 0.0, 0.0), tuning="table"):
        """tuning: "table" takes the committed per-grid tiling
        (inq_amd.tuning) for this grid length and max_bands, None keeps the
        library's defaults, a dict sets the knobs of inqamd_tuning."""
        self.system = system
        self.tables = system.tables
        self.n = tuple(system.n)
        self._h = _Handle(system.n, system.lattice, system.vloc, system.tables, max_bands)
        self._A = np.zeros(3)
        self._k = np.zeros(3)
        self.kpoint = kpoint
        self._commutator = "inq"
        if tuning == "table":
            from . import tuning as _tuning
            tuning = _tuning.lookup(self.n, max_bands)
        if tuning:
            self.tuning = tuning

    @property
    def tuning(self):
        """The per-handle tiling knobs (inqamd_tuning; 0 = library default)."""
        t = _lib.Tuning()
        _lib.check(self._h._lib.inqamd_get_tuning(self._h.h, ctypes.byref(t)), self._h.h)
        return {f: getattr(t, f) for f, _ in _lib.Tuning._fields_}

    @tuning.setter
    def tuning(self, knobs):
        t = _lib.Tuning()
        for f, v in dict(knobs).items():
            if f not in dict(_lib.Tuning._fields_):
                raise ValueError(f"unknown tuning knob {f!r}")
            setattr(t, f, int(v))
        _lib.check(self._h._lib.inqamd_set_tuning(self._h.h, ctypes.byref(t)), self._h.h)

    @property
    def handle(self):
        return self._h

    @property
    def max_bands(self):
        return self._h.max_bands

    @property
    def uniform_vector_potential(self):
        return self._A.copy()

    @uniform_vector_potential.setter
    def uniform_vector_potential(self, A_cov):
        self._A = np.asarray(A_cov, float).reshape(3).copy()
        self._h.call("inqamd_set_uniform_vector_potential", _lib.dvec(self._A))

    @property
    def kpoint(self):
        return self._k.copy()

    @kpoint.setter
    def kpoint(self, k_cov):
        self._k = np.asarray(k_cov, float).reshape(3).copy()
        self._h.call("inqamd_set_kpoint", _lib.dvec(self._k))

    @property
    def position_commutator(self):
        """The current's [V_nl, r] form: "inq" (default; projector_all.hpp:
        319-403 as written, the plain projection scaled by the KB coefficient
        twice) or "corrected" (each projection scaled once: J = -dE/dA)."""
        return self._commutator

    @position_commutator.setter
    def position_commutator(self, name):
        self._h.call("inqamd_set_position_commutator", _lib.commutator_code(name))
        self._commutator = name

    def set_scalar_potential(self, vloc):
        v = vloc.to(dtype=torch.float64).contiguous()
        self._h.call("inqamd_set_scalar_potential", ctypes.c_void_p(v.data_ptr()))
        self._vkeep = v

    def scalar_potential(self, device=None):
        """The scalar potential last given to the Hamiltonian from the host side
        (the system's V, or the last set_scalar_potential), as a float64 device
        tensor [N] (ks_hamiltonian.hpp:200-202).  A potential the library
        updates itself (tddft.ks_potential) is not reflected here."""
        v = getattr(self, "_vkeep", None)
        if v is None:
            dev = device or torch.device("cuda", torch.cuda.current_device())
            v = torch.from_numpy(np.ascontiguousarray(np.ravel(self.system.vloc), np.float64)).to(dev)
            self._vkeep = v
        return v

    def __call__(self, phi, out=None, fourier=False):
        """H phi for real-space orbitals (ks_hamiltonian.hpp:132-152) or, with
        fourier=True, for G-space orbitals (:156-176)."""
        p, nst, ld = _orbital_args(phi)
        if out is None:
            out = torch.empty_like(phi)
        q, _, _ = _orbital_args(out)
        name = "inqamd_hamiltonian_apply_fourier" if fourier else "inqamd_hamiltonian_apply"
        self._h.call(name, p, q, nst, ld)
        return out

    def profile(self, phi, out):
        """One Hpsi into `out` with per-kernel HIP event timing:
        {name: (ms, launches)} for the projector and the three sweeps."""
        p, nst, ld = _orbital_args(phi)
        q, _, _ = _orbital_args(out)
        ms = (ctypes.c_double * 4)()
        nl = (ctypes.c_int * 4)()
        self._h.call("inqamd_hamiltonian_profile", p, q, nst, ld, ms, nl)
        names = ("projector", "z_fwd", "xy_kin", "z_bwd_epi")
        return {n: (ms[i], nl[i]) for i, n in enumerate(names)}

    def non_local(self, phi):
        p, nst, ld = _orbital_args(phi)
        out = torch.empty_like(phi)
        self._h.call("inqamd_non_local", p, ctypes.c_void_p(out.data_ptr()), nst, ld)
        return out

    def projectors_all(self):
        return projector_all(self)

    def band_energies(self, hphi, phi):
        """energy::calculate eigenvalues (energy.hpp:74-75)."""
        p, nst, ld = _orbital_args(phi)
        q, _, _ = _orbital_args(hphi)
        eps = torch.empty(nst, dtype=torch.float64, device=phi.device)
        self._h.call("inqamd_band_energies", q, p, ctypes.c_void_p(eps.data_ptr()), nst, ld)
        return eps

    def energy(self, phi, occ=None):
        """The per-band terms of energy::calculate (energy.hpp:59-87).

        Returns device arrays (eigenvalues, normres, nl_me) of length nst:
        eps_b = Re<H phi_b|phi_b>/<phi_b|phi_b>, normres_b = dV |H phi_b -
        eps_b phi_b|^2 and nl_b = Re<V_nl phi_b|phi_b>/<phi_b|phi_b>.  With
        ``occ`` (a sequence of nst occupations) it returns instead the
        occupation-weighted sums (eigenvalues, non_local) the reference
        accumulates with occ_sum, plus the normres array.
        """
        p, nst, ld = _orbital_args(phi)
        out = torch.empty(3, nst, dtype=torch.float64, device=phi.device)
        self._h.call("inqamd_energy", p, nst, ld, ctypes.c_void_p(out[0].data_ptr()),
                     ctypes.c_void_p(out[1].data_ptr()), ctypes.c_void_p(out[2].data_ptr()))
        eps, normres, nl = out[0], out[1], out[2]
        if occ is None:
            return eps, normres, nl
        w = torch.as_tensor(occ, dtype=torch.float64, device=phi.device)
        if w.shape != (nst,):
            raise ValueError("occ must hold one occupation per band")
        return float((w * eps).sum()), float((w * nl).sum()), normres

    def current(self, phi, occ=2.0):
        """observables::current (current.hpp:25-53) for bands of equal
        occupation: (J_cov, J_cart) as numpy vectors (synchronises)."""
        p, nst, ld = _orbital_args(phi)
        jc = (ctypes.c_double * 3)()
        jx = (ctypes.c_double * 3)()
        self._h.call("inqamd_current", p, nst, ld, ctypes.c_double(occ), jc, jx)
        return np.array(list(jc)), np.array(list(jx))

    def close(self):
        self._h.close()
