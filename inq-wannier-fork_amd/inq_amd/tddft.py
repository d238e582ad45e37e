"""Interacting real-time TDDFT for static ions and one spin component.

Mirrors observables::density::calculate (src/observables/density.hpp:23-95),
solvers::poisson (src/solvers/poisson.hpp, 3-D periodic), the LDA of
options::theory::lda (LDA_X + LDA_C_PZ through hamiltonian::xc_term),
self_consistency::update_hamiltonian (src/hamiltonian/self_consistency.hpp:
88-162) and real_time::etrs with its self-consistent second half
(src/real_time/etrs.hpp:22-82).  Fields are float64 device tensors shaped
like the grid; everything runs in libinq_amd.so on the GPU.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .hamiltonian import _orbital_args
from .real_time import _to_covariant


def _field(t, ham):
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64
            and t.is_contiguous() and t.numel() == int(np.prod(ham.n))):
        raise ValueError("fields must be contiguous float64 GPU tensors with one value per point")
    return ctypes.c_void_p(t.data_ptr())


def _occ(occ, nst):
    o = np.ascontiguousarray(np.broadcast_to(np.asarray(occ, np.float64), (nst,)))
    return o, o.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def density(ham, psi, occ=2.0):
    """sum_b occ_b |psi_b|^2 on the grid (occ: scalar or one value per band)."""
    p, nst, ld = _orbital_args(psi)
    rho = torch.empty(ham.n, dtype=torch.float64, device=psi.device)
    o, op = _occ(occ, nst)
    ham.handle.call("inqamd_density", p, nst, ld, op, _field(rho, ham))
    return rho


def dipole(ham, rho):
    """observables::dipole of a density (cartesian, the electronic part)."""
    out = np.zeros(3)
    ham.handle.call("inqamd_dipole", _field(rho, ham), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def poisson(ham, rho):
    """The periodic Hartree potential of rho (zero mean)."""
    vh = torch.empty_like(rho)
    ham.handle.call("inqamd_poisson", _field(rho, ham), _field(vh, ham))
    return vh


def lda_xc(ham, rho):
    """LDA exchange-correlation: (vxc, exc, nvxc)."""
    vxc = torch.empty_like(rho)
    exc, nvxc = ctypes.c_double(0.0), ctypes.c_double(0.0)
    ham.handle.call("inqamd_lda_xc", _field(rho, ham), _field(vxc, ham), ctypes.byref(exc),
                    ctypes.byref(nvxc))
    return vxc, exc.value, nvxc.value


XC_KINDS = {"lda_x": 1, "lda": 2, "pbe_x": 3, "pbe_c": 4, "pbe": 5}


def xc(ham, rho, kind="pbe"):
    """One exchange-correlation functional (lda_x, lda, pbe_x, pbe_c, pbe):
    (vxc, exc, nvxc)."""
    vxc = torch.empty_like(rho)
    exc, nvxc = ctypes.c_double(0.0), ctypes.c_double(0.0)
    ham.handle.call("inqamd_xc", _field(rho, ham), XC_KINDS[kind], _field(vxc, ham),
                    ctypes.byref(exc), ctypes.byref(nvxc))
    return vxc, exc.value, nvxc.value


def _xc_code(xc):
    """ks_potential's theory: False/None -> 0, True/'lda' -> 1, 'pbe' -> 2."""
    if xc is None or xc is False:
        return 0
    if xc is True or xc == "lda":
        return 1
    if xc == "pbe":
        return 2
    raise ValueError("xc must be False, True/'lda' or 'pbe'")


def ks_potential(ham, vion, rho, hartree=True, xc=True):
    """update_hamiltonian: the Hamiltonian's scalar potential becomes
    vion + V_H[rho] + V_xc[rho] (xc: False, True/'lda' or 'pbe'); returns
    {external, hartree, exc, nvxc}."""
    en = (ctypes.c_double * 4)()
    ham.handle.call("inqamd_ks_potential", _field(vion, ham), _field(rho, ham), int(hartree),
                    _xc_code(xc), en)
    return dict(zip(("external", "hartree", "exc", "nvxc"), list(en)))


def etrs(ham, psi, vion, time, dt, occ=2.0, pert=None, order=4, hartree=True, xc=True):
    """One interacting ETRS step; the Hamiltonian's scalar potential must be
    V_KS of H(t) (ks_potential of the current density) and is V_KS of
    H(t+dt) afterwards.  Returns the SCF passes of the second half."""
    p, nst, ld = _orbital_args(psi)
    lat = ham.system.lattice
    if pert is not None and pert.has_uniform_vector_potential():
        A_t = _to_covariant(lat, pert.uniform_vector_potential(time))
        A_tdt = _to_covariant(lat, pert.uniform_vector_potential(time + dt))
    else:
        A_t = np.zeros(3)
        A_tdt = np.zeros(3)
    o, op = _occ(occ, nst)
    nscf = ctypes.c_int(0)
    ham.handle.call("inqamd_etrs_scf_step", ctypes.c_double(dt), _lib.dvec(A_t), _lib.dvec(A_tdt),
                    int(order), p, nst, ld, _field(vion, ham), op, int(hartree), _xc_code(xc),
                    ctypes.byref(nscf))
    return nscf.value
