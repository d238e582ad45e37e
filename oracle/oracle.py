"""ctypes wrapper of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
import this module.  It loads ``oracle/_build/liboracle.so`` (built by
``oracle/Makefile``) and exposes numpy-level helpers.  Arrays are numpy
complex128 in INQ's [x][y][z][band] layout.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)
_c_flt_p = ctypes.POINTER(ctypes.c_float)


class OrcSystem(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int * 3),
        ("lattice", ctypes.c_double * 9),
        ("vloc", _c_dbl_p),
        ("nproj", ctypes.c_int),
        ("max_sphere", ctypes.c_int),
        ("max_nlm", ctypes.c_int),
        ("points", _c_int_p),
        ("positions", _c_flt_p),
        ("matrices", _c_dbl_p),
        ("coeff", _c_dbl_p),
        ("nlm", _c_int_p),
        ("sphere_size", _c_int_p),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library not built: {_LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_fft3d.argtypes = [_c_int_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.orc_fft1d.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.orc_containing_cube.argtypes = [_c_dbl_p, _c_int_p, _c_dbl_p, ctypes.c_double,
                                          _c_int_p, _c_int_p]
        L.orc_sphere_build.argtypes = [_c_dbl_p, _c_int_p, _c_dbl_p, ctypes.c_double, ctypes.c_int,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, _c_int_p]
        L.orc_gaussian_projector_matrix.argtypes = [_c_dbl_p, ctypes.c_int, ctypes.c_void_p,
                                                    ctypes.c_void_p, ctypes.c_double, ctypes.c_int,
                                                    ctypes.c_void_p, ctypes.c_int]
        sysp = ctypes.POINTER(OrcSystem)
        L.orc_hamiltonian_apply.argtypes = [sysp, _c_dbl_p, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_void_p, ctypes.c_void_p]
        L.orc_hamiltonian_apply_fourier.argtypes = [sysp, _c_dbl_p, ctypes.c_int, ctypes.c_int,
                                                    ctypes.c_void_p, ctypes.c_void_p]
        L.orc_energy.argtypes = [sysp, _c_dbl_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_projector_project.argtypes = [sysp, _c_dbl_p, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_void_p, ctypes.c_void_p]
        L.orc_projector_apply.argtypes = [sysp, _c_dbl_p, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_void_p, ctypes.c_void_p]
        L.orc_band_energies.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p]
        L.orc_exponential_in_place.argtypes = [sysp, _c_dbl_p, ctypes.c_double, ctypes.c_double,
                                               ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_void_p]
        L.orc_exponential_2_for_1.argtypes = [sysp, _c_dbl_p, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                              ctypes.c_void_p]
        L.orc_etrs_step.argtypes = [sysp, _c_dbl_p, _c_dbl_p, ctypes.c_double, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.orc_randomize.argtypes = [_c_int_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.orc_current.argtypes = [sysp, _c_dbl_p, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_void_p, ctypes.c_int, _c_dbl_p, _c_dbl_p]
        L.orc_td_current.argtypes = [sysp, ctypes.c_int, _c_dbl_p, ctypes.c_double, _c_dbl_p,
                                     ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_double)]
        L.orc_laplacian_add.argtypes = [sysp, ctypes.c_double, _c_dbl_p, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_preconditioner.argtypes = [sysp, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.orc_orthogonalize.argtypes = [ctypes.c_long, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_void_p]
        L.orc_steepest_descent.argtypes = [sysp, _c_dbl_p, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_void_p]
        L.orc_steepest_descent_dense.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                                 ctypes.c_double, ctypes.c_int, ctypes.c_void_p,
                                                 ctypes.c_void_p]
        L.orc_density.argtypes = [ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p]
        L.orc_poisson.argtypes = [sysp, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_lda_xc.argtypes = [ctypes.c_long, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.orc_dipole.argtypes = [sysp, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_xc.argtypes = [sysp, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                             ctypes.POINTER(ctypes.c_double)]
        L.orc_ks_potential.argtypes = [sysp, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_etrs_scf.argtypes = [sysp, ctypes.c_void_p, ctypes.c_void_p, _c_dbl_p, _c_dbl_p,
                                   ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.POINTER(ctypes.c_int)]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _ivec(v):
    return (ctypes.c_int * len(v))(*[int(x) for x in v])


def _dvec(v):
    return (ctypes.c_double * len(v))(*[float(x) for x in v])


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with status {rc}")


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def _as_points_bands(a):
    """View an [nx][ny][nz][nst] (or [N][nst]) complex array as contiguous."""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a


def fft3d(a, sign=-1, normalize=False):
    """Batched 3-D DFT over the first three axes of a [nx][ny][nz][nst] array."""
    a = _as_points_bands(a)
    n = a.shape[:3]
    nst = a.shape[3]
    out = np.empty_like(a)
    _check(lib().orc_fft3d(_ivec(n), nst, nst, _ptr(a), _ptr(out), int(sign), int(normalize)),
           "fft3d")
    return out


def to_fourier(a):
    return fft3d(a, -1, False)


def to_real(a, normalize=True):
    return fft3d(a, +1, normalize)


def fft1d(v, sign=-1):
    v = np.ascontiguousarray(v, dtype=np.complex128)
    out = np.empty_like(v)
    _check(lib().orc_fft1d(v.size, _ptr(v), _ptr(out), int(sign)), "fft1d")
    return out


def containing_cube(lattice, n, pos, radius):
    lo = (ctypes.c_int * 3)()
    hi = (ctypes.c_int * 3)()
    _check(lib().orc_containing_cube(_dvec(np.ravel(lattice)), _ivec(n), _dvec(pos), float(radius),
                                     lo, hi), "containing_cube")
    return list(lo), list(hi)


def sphere_build(lattice, n, center, radius):
    """Returns (coords int32 [K,3], distance float32 [K], relpos float32 [K,3])."""
    cap = 4096
    while True:
        coords = np.zeros((cap, 3), np.int32)
        dist = np.zeros(cap, np.float32)
        rel = np.zeros((cap, 3), np.float32)
        cnt = ctypes.c_int(0)
        rc = lib().orc_sphere_build(_dvec(np.ravel(lattice)), _ivec(n), _dvec(center), float(radius),
                                    cap, _ptr(coords), _ptr(dist), _ptr(rel), ctypes.byref(cnt))
        if rc == 2:
            cap = cnt.value
            continue
        _check(rc, "sphere_build")
        k = cnt.value
        return coords[:k].copy(), dist[:k].copy(), rel[:k].copy()


def gaussian_projector_matrix(lattice, distance, relpos, sigma, nlm=4):
    k = distance.size
    m = np.zeros((nlm, k), np.float64)
    d = np.ascontiguousarray(distance, np.float32)
    r = np.ascontiguousarray(relpos, np.float32)
    _check(lib().orc_gaussian_projector_matrix(_dvec(np.ravel(lattice)), k, _ptr(d), _ptr(r),
                                               float(sigma), nlm, _ptr(m), k), "projector matrix")
    return m


class System:
    """Holds numpy arrays alive and the OrcSystem struct pointing at them."""

    def __init__(self, n, lattice, vloc, tables=None):
        self.n = tuple(int(x) for x in n)
        self.lattice = np.ascontiguousarray(np.reshape(lattice, (3, 3)), np.float64)
        self.vloc = np.ascontiguousarray(np.ravel(vloc), np.float64)
        s = OrcSystem()
        s.n = (ctypes.c_int * 3)(*self.n)
        s.lattice = (ctypes.c_double * 9)(*self.lattice.ravel())
        s.vloc = self.vloc.ctypes.data_as(_c_dbl_p)
        if tables is None or tables["nproj"] == 0:
            self.keep = []
            s.nproj = 0
            s.max_sphere = 0
            s.max_nlm = 0
        else:
            pts = np.ascontiguousarray(tables["points"], np.int32)
            pos = np.ascontiguousarray(tables["positions"], np.float32)
            mat = np.ascontiguousarray(tables["matrices"], np.float64)
            coe = np.ascontiguousarray(tables["coeff"], np.float64)
            nlm = np.ascontiguousarray(tables["nlm"], np.int32)
            ssz = np.ascontiguousarray(tables["sphere_size"], np.int32)
            self.keep = [pts, pos, mat, coe, nlm, ssz]
            s.nproj = int(tables["nproj"])
            s.max_sphere = int(tables["max_sphere"])
            s.max_nlm = int(tables["max_nlm"])
            s.points = pts.ctypes.data_as(_c_int_p)
            s.positions = pos.ctypes.data_as(_c_flt_p)
            s.matrices = mat.ctypes.data_as(_c_dbl_p)
            s.coeff = coe.ctypes.data_as(_c_dbl_p)
            s.nlm = nlm.ctypes.data_as(_c_int_p)
            s.sphere_size = ssz.ctypes.data_as(_c_int_p)
        self.s = s

    @property
    def npoints(self):
        return self.n[0] * self.n[1] * self.n[2]

    def hamiltonian(self, psi, kA=(0.0, 0.0, 0.0)):
        psi = np.ascontiguousarray(psi, np.complex128)
        nst = psi.shape[-1]
        out = np.empty_like(psi)
        _check(lib().orc_hamiltonian_apply(ctypes.byref(self.s), _dvec(kA), nst, nst, _ptr(psi),
                                           _ptr(out)), "hamiltonian_apply")
        return out

    def energy(self, psi, kA=(0.0, 0.0, 0.0)):
        """energy::calculate per band: (eigenvalues, residual norms, non-local)."""
        psi = np.ascontiguousarray(psi, np.complex128)
        nst = psi.shape[-1]
        eps, nr, nl = np.zeros(nst), np.zeros(nst), np.zeros(nst)
        _check(lib().orc_energy(ctypes.byref(self.s), _dvec(kA), nst, nst, _ptr(psi), _ptr(eps),
                                _ptr(nr), _ptr(nl)), "energy")
        return eps, nr, nl

    def laplacian_add(self, ff, laplff, factor=1.0, gradcoeff=(0.0, 0.0, 0.0)):
        """operations::laplacian_add on G-space arrays (laplacian.hpp:24-38); returns
        laplff + factor (-|G|^2 + gradcoeff . G) ff (gradcoeff contravariant)."""
        ff = np.ascontiguousarray(ff, np.complex128)
        out = np.array(laplff, dtype=np.complex128, order="C", copy=True)
        nst = ff.shape[-1]
        _check(lib().orc_laplacian_add(ctypes.byref(self.s), float(factor), _dvec(gradcoeff), nst,
                                       nst, _ptr(ff), _ptr(out)), "laplacian_add")
        return out

    def preconditioner(self, phi):
        """operations::preconditioner on G-space orbitals (preconditioner.hpp:25-55)."""
        out = np.array(phi, dtype=np.complex128, order="C", copy=True)
        nst = out.shape[-1]
        _check(lib().orc_preconditioner(ctypes.byref(self.s), nst, nst, _ptr(out)),
               "preconditioner")
        return out

    def steepest_descent(self, phi, kA=(0.0, 0.0, 0.0), precondition=True):
        """eigensolvers::steepest_descent on G-space orbitals (steepest_descent.hpp:23-88)
        with the G-space Hamiltonian; returns the new orbitals."""
        out = np.array(phi, dtype=np.complex128, order="C", copy=True)
        nst = out.shape[-1]
        _check(lib().orc_steepest_descent(ctypes.byref(self.s), _dvec(kA), int(bool(precondition)),
                                          nst, nst, _ptr(out)), "steepest_descent")
        return out

    def subspace_diagonalization(self, phi, kA=(0.0, 0.0, 0.0)):
        """ground_state::subspace_diagonalization (subspace_diagonalization.hpp:16-25)
        on real-space orbitals: H_sub = dV Phi^H (H Phi) (overlap.hpp:26-60), its
        eigenpairs (numpy eigh, ascending), Phi <- Phi V.  Returns (eigenvalues, phi)."""
        phi = np.ascontiguousarray(phi, np.complex128)
        nst = phi.shape[-1]
        N = int(np.prod(self.n))
        dV = abs(np.linalg.det(np.asarray(self.lattice, float))) / N
        hphi = self.hamiltonian(phi, kA).reshape(N, nst)
        m = phi.reshape(N, nst)
        hs = dV * (m.conj().T @ hphi)
        w, v = np.linalg.eigh(0.5 * (hs + hs.conj().T))
        return w, (m @ v).reshape(phi.shape)

    def poisson(self, rho):
        """solvers::poisson, 3-D periodic, of a real density [nx][ny][nz]."""
        rho = np.ascontiguousarray(rho, np.float64)
        vh = np.zeros_like(rho)
        _check(lib().orc_poisson(ctypes.byref(self.s), _ptr(rho), _ptr(vh)), "poisson")
        return vh

    def dipole(self, rho):
        """observables::dipole of a density: dV sum r_cart rho."""
        rho = np.ascontiguousarray(rho, np.float64)
        out = np.zeros(3)
        _check(lib().orc_dipole(ctypes.byref(self.s), _ptr(rho), _ptr(out)), "dipole")
        return out

    XC_KINDS = {"lda_x": 1, "lda": 2, "pbe_x": 3, "pbe_c": 4, "pbe": 5}

    def xc(self, rho, kind):
        """One exchange-correlation functional of a density (xc_term.hpp:254-300):
        (vxc, exc) for kind in lda_x, lda (LDA_X + LDA_C_PZ), pbe_x, pbe_c, pbe."""
        rho = np.ascontiguousarray(rho, np.float64)
        vxc = np.zeros_like(rho)
        exc = ctypes.c_double(0.0)
        _check(lib().orc_xc(ctypes.byref(self.s), _ptr(rho), self.XC_KINDS[kind], _ptr(vxc),
                            ctypes.byref(exc)), "xc")
        return vxc, exc.value

    def ks_potential(self, vion, rho, hartree=True, xc=True):
        """self_consistency::update_hamiltonian: (vks, {external, hartree, exc, nvxc})."""
        vion = np.ascontiguousarray(vion, np.float64)
        rho = np.ascontiguousarray(rho, np.float64)
        vks = np.zeros_like(rho)
        en = np.zeros(4)
        _check(lib().orc_ks_potential(ctypes.byref(self.s), _ptr(vion), _ptr(rho), int(hartree),
                                      _xc_code(xc), _ptr(vks), _ptr(en)), "ks_potential")
        return vks, dict(zip(("external", "hartree", "exc", "nvxc"), en))

    def etrs_scf(self, psi, vion, occ, dt, kA_t=(0.0, 0.0, 0.0), kA_tdt=(0.0, 0.0, 0.0), order=4,
                 hartree=True, xc=True):
        """real_time::etrs with its SCF re-check; this System's vloc is V_KS(t).
        Returns (psi(t+dt), V_KS(t+dt), scf passes)."""
        out = np.array(psi, dtype=np.complex128, order="C", copy=True)
        nst = out.shape[-1]
        vion = np.ascontiguousarray(vion, np.float64)
        occ = np.ascontiguousarray(np.broadcast_to(occ, (nst,)), np.float64)
        vks = np.zeros(self.n)
        nscf = ctypes.c_int(0)
        _check(lib().orc_etrs_scf(ctypes.byref(self.s), _ptr(vion), _ptr(occ), _dvec(kA_t),
                                  _dvec(kA_tdt), float(dt), int(order), int(hartree), _xc_code(xc), nst,
                                  nst, _ptr(out), _ptr(vks), ctypes.byref(nscf)), "etrs_scf")
        return out, vks, nscf.value

    def hamiltonian_fourier(self, phi, kA=(0.0, 0.0, 0.0)):
        """ks_hamiltonian::operator() on G-space orbitals (ks_hamiltonian.hpp:156-176)."""
        phi = np.ascontiguousarray(phi, np.complex128)
        nst = phi.shape[-1]
        out = np.zeros_like(phi)
        _check(lib().orc_hamiltonian_apply_fourier(ctypes.byref(self.s), _dvec(kA), nst, nst,
                                                   _ptr(phi), _ptr(out)), "hamiltonian_fourier")
        return out

    def project(self, psi, kA=(0.0, 0.0, 0.0)):
        psi = np.ascontiguousarray(psi, np.complex128)
        nst = psi.shape[-1]
        out = np.zeros((self.s.nproj, self.s.max_sphere, nst), np.complex128)
        _check(lib().orc_projector_project(ctypes.byref(self.s), _dvec(kA), nst, nst, _ptr(psi),
                                           _ptr(out)), "projector_project")
        return out

    def apply(self, sphere, hpsi, kA=(0.0, 0.0, 0.0)):
        sphere = np.ascontiguousarray(sphere, np.complex128)
        hpsi = np.array(hpsi, dtype=np.complex128, order="C", copy=True)
        nst = hpsi.shape[-1]
        _check(lib().orc_projector_apply(ctypes.byref(self.s), _dvec(kA), nst, nst, _ptr(sphere),
                                         _ptr(hpsi)), "projector_apply")
        return hpsi

    def exponential_in_place(self, psi, z, order=4, kA=(0.0, 0.0, 0.0)):
        psi = np.array(psi, dtype=np.complex128, order="C", copy=True)
        nst = psi.shape[-1]
        _check(lib().orc_exponential_in_place(ctypes.byref(self.s), _dvec(kA), float(np.real(z)),
                                              float(np.imag(z)), int(order), nst, nst, _ptr(psi)),
               "exponential_in_place")
        return psi

    def exponential_2_for_1(self, psi, z1, z2, order=4, kA=(0.0, 0.0, 0.0)):
        phi = np.array(psi, dtype=np.complex128, order="C", copy=True)
        exp = np.empty_like(phi)
        nst = phi.shape[-1]
        _check(lib().orc_exponential_2_for_1(ctypes.byref(self.s), _dvec(kA), float(np.real(z1)),
                                             float(np.imag(z1)), float(np.real(z2)),
                                             float(np.imag(z2)), int(order), nst, nst, _ptr(phi),
                                             _ptr(exp)), "exponential_2_for_1")
        return phi, exp

    def etrs_step(self, psi, dt, kA_t, kA_tdt, order=4):
        psi = np.array(psi, dtype=np.complex128, order="C", copy=True)
        nst = psi.shape[-1]
        _check(lib().orc_etrs_step(ctypes.byref(self.s), _dvec(kA_t), _dvec(kA_tdt), float(dt),
                                   int(order), nst, nst, _ptr(psi)), "etrs_step")
        return psi


def randomize(n, nst, index=0):
    """operations::randomize for complex orbitals: [nx, ny, nz, nst]."""
    out = np.zeros(tuple(n) + (nst,), np.complex128)
    _check(lib().orc_randomize(_ivec(n), int(nst), int(index), _ptr(out)), "randomize")
    return out


COMMUTATORS = {"inq": 0, "corrected": 1}


def _commutator(name):
    if name not in COMMUTATORS:
        raise ValueError(f"commutator must be one of {sorted(COMMUTATORS)}, not {name!r}")
    return COMMUTATORS[name]


def _system_current(self, psi, kA=(0.0, 0.0, 0.0), occ=2.0, commutator="inq"):
    """observables::current; commutator "inq" is position_commutator as the
    reference writes it, "corrected" scales both projections once."""
    psi = np.ascontiguousarray(psi, np.complex128)
    nst = psi.shape[-1]
    jc = (ctypes.c_double * 3)()
    jx = (ctypes.c_double * 3)()
    _check(lib().orc_current(ctypes.byref(self.s), _dvec(kA), float(occ), nst, nst, _ptr(psi),
                             _commutator(commutator), jc, jx), "current")
    return np.array(list(jc)), np.array(list(jx))


def _system_td_current(self, psi, coef, omega, direction, dt, nsteps, order=4, occ=2.0,
                       commutator="inq"):
    psi = np.array(psi, dtype=np.complex128, order="C", copy=True)
    nst = psi.shape[-1]
    c = np.ascontiguousarray(coef, np.float64)
    loss = ctypes.c_double(0.0)
    _check(lib().orc_td_current(ctypes.byref(self.s), c.size, c.ctypes.data_as(_c_dbl_p),
                                float(omega), _dvec(direction), float(dt), int(nsteps), int(order),
                                float(occ), nst, nst, _ptr(psi), _commutator(commutator),
                                ctypes.byref(loss)), "td_current")
    return loss.value, psi


System.current = _system_current
System.td_current = _system_td_current


def band_energies(hpsi, psi):
    hpsi = np.ascontiguousarray(hpsi, np.complex128)
    psi = np.ascontiguousarray(psi, np.complex128)
    nst = psi.shape[-1]
    npoint = psi.size // nst
    eps = np.zeros(nst, np.float64)
    _check(lib().orc_band_energies(npoint, nst, nst, _ptr(hpsi), _ptr(psi), _ptr(eps)),
           "band_energies")
    return eps


def orthogonalize(phi, dV):
    """operations::orthogonalize (orthogonalize.hpp:25-32) with volume element dV."""
    out = np.array(phi, dtype=np.complex128, order="C", copy=True)
    nst = out.shape[-1]
    npoint = out.size // nst
    _check(lib().orc_orthogonalize(npoint, nst, nst, float(dV), _ptr(out)), "orthogonalize")
    return out


def steepest_descent_dense(matrix, phi, dV, niter):
    """niter eigensolvers::steepest_descent calls with a dense matrix operator and no
    preconditioner (the reference's unit-test setting); returns (phi, eigenvalues)."""
    M = np.ascontiguousarray(matrix, np.complex128)
    out = np.array(phi, dtype=np.complex128, order="C", copy=True)
    npoint, nvec = out.shape
    ev = np.zeros(nvec, np.complex128)
    _check(lib().orc_steepest_descent_dense(npoint, nvec, _ptr(M), float(dV), int(niter),
                                            _ptr(out), _ptr(ev)), "steepest_descent_dense")
    return out, ev


def _xc_code(xc):
    """The theory of ks_potential/etrs_scf: False -> 0, True/'lda' -> 1, 'pbe' -> 2."""
    if xc is None or xc is False:
        return 0
    if xc is True or xc == "lda":
        return 1
    if xc == "pbe":
        return 2
    raise ValueError("xc must be False, True/'lda' or 'pbe'")


def density(psi, occ):
    """observables::density::calculate_add for one spin: sum_b occ_b |psi_b|^2."""
    psi = np.ascontiguousarray(psi, np.complex128)
    nst = psi.shape[-1]
    npoint = psi.size // nst
    occ = np.ascontiguousarray(np.broadcast_to(occ, (nst,)), np.float64)
    rho = np.zeros(psi.shape[:-1])
    _check(lib().orc_density(npoint, nst, nst, _ptr(occ), _ptr(psi), _ptr(rho)), "density")
    return rho


def lda_xc(rho, dV):
    """LDA_X + LDA_C_PZ through xc_term: (vxc, exc, nvxc)."""
    rho = np.ascontiguousarray(rho, np.float64)
    vxc = np.zeros_like(rho)
    exc, nvxc = ctypes.c_double(0.0), ctypes.c_double(0.0)
    _check(lib().orc_lda_xc(rho.size, float(dV), _ptr(rho), _ptr(vxc), ctypes.byref(exc),
                            ctypes.byref(nvxc)), "lda_xc")
    return vxc, exc.value, nvxc.value
