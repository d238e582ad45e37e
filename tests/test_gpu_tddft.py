"""GPU tests of the interacting TDDFT pieces against the oracle: density,
Poisson (plus the reference's point-charge known answers, through the GPU's
prime-length FFT), LDA exchange-correlation, the Kohn-Sham potential update
and the self-consistent ETRS step on a pencil-FFT and a three-sweep grid.
Tolerance 1e-12 relative unless stated."""
import json
import os
import sys

import numpy as np
import pytest
import torch

import oracle as O
import inq_amd
from inq_amd import systems, tddft
from helpers import relerr

pytestmark = pytest.mark.gpu
TOL = 1e-12
HERE = os.path.dirname(os.path.abspath(__file__))
KNOWN = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def to_dev(a, device, dtype=np.complex128):
    return torch.from_numpy(np.ascontiguousarray(a, dtype)).to(device)


@pytest.fixture(scope="module")
def cfg(device):
    s = systems.diamond_system(1)
    psi = systems.random_orbitals(s, 8)
    return s, psi, O.System(s.n, s.lattice, s.vloc, s.tables)


def test_density_matches_oracle(device, cfg):
    s, psi, _ = cfg
    say("density")
    occ = np.linspace(2.0, 0.5, 8)
    ham = inq_amd.ks_hamiltonian(s, max_bands=8)
    rho = tddft.density(ham, to_dev(psi, device), occ).cpu().numpy()
    assert relerr(rho, O.density(psi, occ)) < TOL
    ham.close()


def test_poisson_point_charge_reference_gpu(device):
    k = KNOWN["poisson_point_charge"]
    n = tuple(k["grid"])
    say("poisson point charge", n)
    s = systems.make_system(n, np.diag(k["cell"]), np.zeros((0, 3)), seed=0, system_id=91,
                            with_projectors=False)
    ham = inq_amd.ks_hamiltonian(s, max_bands=1)
    rho = torch.zeros(n, dtype=torch.float64, device=device)
    rho[0, 0, 0] = k["charge"]
    v = tddft.poisson(ham, rho).cpu().numpy()
    say("sum |V| %.10f, V(0) %.10f" % (np.abs(v).sum(), v[0, 0, 0]))
    assert np.abs(v).sum() == pytest.approx(k["sum_abs_potential"], abs=5e-10)
    assert v[0, 0, 0] == pytest.approx(k["potential_at_origin"], abs=5e-11)
    ham.close()


def test_poisson_and_lda_match_oracle(device, cfg):
    s, psi, osys = cfg
    say("poisson + lda")
    ham = inq_amd.ks_hamiltonian(s, max_bands=8)
    rho_h = O.density(psi, 2.0)
    rho = to_dev(rho_h, device, np.float64)
    assert relerr(tddft.poisson(ham, rho).cpu().numpy(), osys.poisson(rho_h)) < TOL
    vxc, exc, nvxc = tddft.lda_xc(ham, rho)
    dV = s.volume / np.prod(s.n)
    rv, re, rn = O.lda_xc(rho_h, dV)
    assert relerr(vxc.cpu().numpy(), rv) < TOL
    assert exc == pytest.approx(re, rel=TOL) and nvxc == pytest.approx(rn, rel=TOL)
    ham.close()


def test_ks_potential_matches_oracle(device, cfg):
    """The update sets the Hamiltonian's potential: Hpsi with it equals the
    oracle's Hpsi with the oracle's V_KS."""
    s, psi, osys = cfg
    say("ks potential")
    ham = inq_amd.ks_hamiltonian(s, max_bands=8)
    rho_h = O.density(psi, 2.0)
    en = tddft.ks_potential(ham, to_dev(s.vloc, device, np.float64),
                            to_dev(rho_h, device, np.float64))
    vks, ren = osys.ks_potential(s.vloc, rho_h)
    for key in ("external", "hartree", "exc", "nvxc"):
        assert en[key] == pytest.approx(ren[key], rel=TOL), key
    h = ham(to_dev(psi, device)).cpu().numpy()
    ref = O.System(s.n, s.lattice, vks, s.tables).hamiltonian(psi)
    assert relerr(h, ref) < TOL
    ham.close()


@pytest.mark.parametrize("grid", ["pencil", "cubic"])
def test_etrs_scf_matches_oracle(device, grid):
    if grid == "pencil":
        lat = np.diag([6.5, 6.5, 6.5])
        s = systems.make_system((12, 14, 12), lat, np.array([[0.3, 0.2, 0.1], [3.4, 3.0, 3.5]]),
                                seed=3, system_id=55)
        nst, nsteps = 4, 3
    else:
        s = systems.diamond_system(1)
        nst, nsteps = 8, 2
    say("etrs scf", grid, s.n)
    psi = systems.random_orbitals(s, nst, ecut=6.0)
    occ = np.full(nst, 2.0)
    coef = systems.rng(0, 0, 4).uniform(-0.05, 0.05, size=8)
    pert = inq_amd.real_time.fourier_series_vector_potential(coef, 0.1)
    dt = 0.02
    ham = inq_amd.ks_hamiltonian(s, max_bands=nst)
    d = to_dev(psi, device)
    vion = to_dev(s.vloc, device, np.float64)
    tddft.ks_potential(ham, vion, tddft.density(ham, d, occ))
    v, _ = O.System(s.n, s.lattice, s.vloc, s.tables).ks_potential(s.vloc, O.density(psi, occ))
    x = psi
    lat = s.lattice
    for it in range(nsteps):
        t = it * dt
        n_gpu = tddft.etrs(ham, d, vion, t, dt, occ, pert)
        A0 = lat @ pert.uniform_vector_potential(t)
        A1 = lat @ pert.uniform_vector_potential(t + dt)
        x, v, n_ref = O.System(s.n, s.lattice, v, s.tables).etrs_scf(x, s.vloc, occ, dt, A0, A1)
        say("step", it, "scf passes gpu", n_gpu, "oracle", n_ref)
        assert n_gpu == n_ref
    err = relerr(d.cpu().numpy(), x)
    say("etrs scf rel err vs oracle %.2e" % err)
    assert err < 1e-11
    # the Hamiltonian now carries V_KS(t + dt): its Hpsi matches the oracle's
    ham.uniform_vector_potential = lat @ pert.uniform_vector_potential(nsteps * dt)
    h = ham(d).cpu().numpy()
    ref = O.System(s.n, s.lattice, v, s.tables).hamiltonian(
        x, lat @ pert.uniform_vector_potential(nsteps * dt))
    assert relerr(h, ref) < 1e-11
    ham.close()


def test_dipole_reference_known_answers_gpu(device):
    k = KNOWN["dipole"]
    lat = np.diag(k["cell"])
    n = tuple(int(round(L / k["spacing"])) for L in k["cell"])
    say("dipole", n)
    s = systems.make_system(n, lat, np.zeros((0, 3)), seed=0, system_id=92, with_projectors=False)
    ham = inq_amd.ks_hamiltonian(s, max_bands=1)
    from helpers import grid_positions
    r = grid_positions(lat, n)
    e = np.exp(-(r * r).sum(-1))
    x, y, z = r[..., 0], r[..., 1], r[..., 2]
    d = tddft.dipole(ham, to_dev(x * e, device, np.float64))
    assert d[0] == pytest.approx(k["x"][0], rel=1e-9) and np.abs(d[1:]).max() < 1e-14
    d = tddft.dipole(ham, to_dev((np.sin(y) + x * z) * e, device, np.float64))
    assert d[0] == pytest.approx(k["xz"][0], abs=1e-10) and d[1] == pytest.approx(k["xz"][1], rel=1e-9)
    rho = to_dev((y + 2.0 * z) * e, device, np.float64)
    assert np.allclose(tddft.dipole(ham, rho), O.System(n, lat, np.zeros(n)).dipole(
        (y + 2.0 * z) * e), rtol=1e-13, atol=1e-14)
    ham.close()


def test_xc_reference_known_answers_gpu(device):
    """LDA_X and GGA_C_PBE on the reference's xc_term density (a 25 x 34 x 11
    orthorhombic grid: spectral gradient and divergence through the 17- and
    11-point axes) reproduce its known answers; every kind matches the
    oracle."""
    k = KNOWN["xc_term"]
    lat = np.diag(k["cell"])
    n = tuple(int(round(L / k["spacing"])) for L in k["cell"])
    from helpers import grid_positions
    r = grid_positions(lat, n)
    f = np.maximum(0.0, np.cos(r @ (2.0 * np.pi / np.array(k["cell"]))) + 1.0)
    rho = f * (k["electrons"] / (f.sum() * np.prod(k["cell"]) / np.prod(n)))
    say("xc known answers", n)
    s = systems.make_system(n, lat, np.zeros((0, 3)), seed=0, system_id=93, with_projectors=False)
    ham = inq_amd.ks_hamiltonian(s, max_bands=1)
    osys = O.System(n, lat, np.zeros(n))
    d = to_dev(rho, device, np.float64)
    for kind in ("lda_x", "lda", "pbe_x", "pbe_c", "pbe"):
        v, exc, nvxc = tddft.xc(ham, d, kind)
        v = v.cpu().numpy()
        vr, er = osys.xc(rho, kind)
        say(kind, "exc %.12f oracle %.12f" % (exc, er))
        assert exc == pytest.approx(er, rel=TOL), kind
        assert np.abs(v - vr).max() < TOL * np.abs(vr).max(), kind
        assert nvxc == pytest.approx((rho * vr).sum() * np.prod(k["cell"]) / np.prod(n), rel=TOL)
        if kind in k:
            vf = v.reshape(-1)
            assert exc == pytest.approx(k[kind]["exc"], abs=1e-9), kind
            assert vf[5439] == pytest.approx(k[kind]["v5439"], abs=1e-10), kind
            assert vf[4444] == pytest.approx(k[kind]["v4444"], abs=1e-10), kind
    ham.close()


def test_pbe_ks_potential_and_etrs_match_oracle(device):
    """The PBE theory through the Kohn-Sham update and two self-consistent
    ETRS steps (three-sweep grid) against the oracle."""
    s = systems.diamond_system(1)
    nst = 8
    psi = systems.random_orbitals(s, nst, ecut=6.0)
    occ = np.full(nst, 2.0)
    say("pbe ks potential + etrs", s.n)
    ham = inq_amd.ks_hamiltonian(s, max_bands=nst)
    d = to_dev(psi, device)
    vion = to_dev(s.vloc, device, np.float64)
    en = tddft.ks_potential(ham, vion, tddft.density(ham, d, occ), xc="pbe")
    v, ren = O.System(s.n, s.lattice, s.vloc, s.tables).ks_potential(s.vloc, O.density(psi, occ),
                                                                     xc="pbe")
    for key in ("external", "hartree", "exc", "nvxc"):
        assert en[key] == pytest.approx(ren[key], rel=TOL), key
    coef = systems.rng(0, 0, 4).uniform(-0.05, 0.05, size=8)
    pert = inq_amd.real_time.fourier_series_vector_potential(coef, 0.1)
    dt, x, lat = 0.02, psi, s.lattice
    for it in range(2):
        t = it * dt
        n_gpu = tddft.etrs(ham, d, vion, t, dt, occ, pert, xc="pbe")
        A0 = lat @ pert.uniform_vector_potential(t)
        A1 = lat @ pert.uniform_vector_potential(t + dt)
        x, v, n_ref = O.System(s.n, s.lattice, v, s.tables).etrs_scf(x, s.vloc, occ, dt, A0, A1,
                                                                     xc="pbe")
        say("step", it, "scf passes gpu", n_gpu, "oracle", n_ref)
        assert n_gpu == n_ref
    assert relerr(d.cpu().numpy(), x) < 1e-11
    ham.close()


def test_etrs_scf_leading_dimension_beyond_max_bands(device):
    """psi as the first nst bands of a wider buffer (ld > max_bands): the
    saved half step is compact, so the step matches the contiguous one and
    leaves the other bands alone."""
    import ctypes
    from inq_amd import _lib
    lat = np.diag([6.5, 6.5, 6.5])
    s = systems.make_system((12, 14, 12), lat, np.array([[0.3, 0.2, 0.1]]), seed=3, system_id=56)
    nst, ld = 4, 8
    psi = systems.random_orbitals(s, nst, ecut=6.0)
    occ = np.full(nst, 2.0)
    vion = to_dev(s.vloc, device, np.float64)
    say("etrs scf ld", ld, "max_bands", nst)
    ha = inq_amd.ks_hamiltonian(s, max_bands=nst)
    a = to_dev(psi, device)
    tddft.ks_potential(ha, vion, tddft.density(ha, a, occ))
    tddft.etrs(ha, a, vion, 0.0, 0.02, occ)
    hb = inq_amd.ks_hamiltonian(s, max_bands=nst)
    wide = np.zeros(psi.shape[:-1] + (ld,), np.complex128)
    wide[..., :nst] = psi
    w = to_dev(wide, device)
    tddft.ks_potential(hb, vion, tddft.density(hb, to_dev(psi, device), occ))
    nscf = ctypes.c_int(0)
    hb.handle.call("inqamd_etrs_scf_step", ctypes.c_double(0.02), _lib.dvec(np.zeros(3)),
                   _lib.dvec(np.zeros(3)), 4, ctypes.c_void_p(w.data_ptr()), nst, ld,
                   ctypes.c_void_p(vion.data_ptr()),
                   np.ascontiguousarray(occ).ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1, 1,
                   ctypes.byref(nscf))
    got = w.cpu().numpy()
    assert relerr(got[..., :nst], a.cpu().numpy()) < 1e-14
    assert np.abs(got[..., nst:]).max() == 0.0
    ha.close()
    hb.close()
