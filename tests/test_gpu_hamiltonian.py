"""GPU parity of the HIP library against the CPU oracle (through the C ABI).

Tolerance: 1e-12 relative (||gpu - oracle||_2 / ||oracle||_2, globally and per
band) for Hpsi, transforms, projector outputs and band energies, FP64.
"""
import ctypes
import sys

import numpy as np
import pytest
import torch

import oracle as O
import inq_amd
from inq_amd import operations, systems
from helpers import band_relerr, grid_positions, relerr

pytestmark = pytest.mark.gpu
TOL = 1e-12


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def to_dev(a, device):
    return torch.from_numpy(np.ascontiguousarray(a, np.complex128)).to(device)


def to_host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.fixture(scope="module")
def cfg0():
    s = systems.diamond_system(1)
    psi = systems.random_orbitals(s, 16)
    return s, psi, O.System(s.n, s.lattice, s.vloc, s.tables)


@pytest.mark.parametrize("n", [(32, 32, 32), (24, 30, 36), (17, 10, 12), (8, 9, 5),
                               (37, 6, 41)])
@pytest.mark.parametrize("nst", [1, 5, 16, 37])
def test_transform_matches_oracle(device, n, nst):
    say("transform", n, nst)
    rng = np.random.default_rng(sum(n) + nst)
    a = rng.standard_normal(n + (nst,)) + 1j * rng.standard_normal(n + (nst,))
    plan = operations.fourier_plan(n, np.diag([5.0, 6.0, 7.0]), max_bands=nst)
    da = to_dev(a, device)
    f = operations.transform.to_fourier(da, plan)
    assert relerr(to_host(f), O.to_fourier(a)) < TOL
    r = operations.space.to_real(f, plan)
    assert relerr(to_host(r), a) < TOL
    # in place
    operations.transform.to_fourier(da, plan, out=da)
    assert relerr(to_host(da), O.to_fourier(a)) < TOL


@pytest.mark.parametrize("kA", [(0.0, 0.0, 0.0), (0.35, -0.21, 0.5)])
def test_hpsi_config0_matches_oracle(device, cfg0, kA):
    s, psi, osys = cfg0
    say("hpsi config0 kA", kA)
    ham = inq_amd.ks_hamiltonian(s, max_bands=16)
    ham.uniform_vector_potential = kA
    hphi = to_host(ham(to_dev(psi, device)))
    ref = osys.hamiltonian(psi, kA)
    err, berr = relerr(hphi, ref), band_relerr(hphi, ref)
    say("hpsi rel err", err, "per-band", berr)
    assert err < TOL and berr < TOL
    eps = ham.band_energies(to_dev(hphi, device), to_dev(psi, device)).cpu().numpy()
    eref = O.band_energies(ref, psi)
    assert np.abs(eps - eref).max() < TOL * np.abs(eref).max()
    ham.close()


def test_hpsi_leading_dimension_and_band_slices(device, cfg0):
    """ld > nst: apply H to the first 5 bands of a 16-band buffer."""
    s, psi, osys = cfg0
    say("hpsi ld slice")
    ham = inq_amd.ks_hamiltonian(s, max_bands=16)
    d = to_dev(psi, device)
    out = torch.zeros_like(d)
    ham.handle.call("inqamd_hamiltonian_apply", ctypes.c_void_p(d.data_ptr()),
                    ctypes.c_void_p(out.data_ptr()), 5, 16)
    got = to_host(out)
    ref = osys.hamiltonian(np.ascontiguousarray(psi[..., :5]))
    assert relerr(got[..., :5], ref) < TOL
    assert np.abs(got[..., 5:]).max() == 0.0
    ham.close()


def test_projector_project_apply_and_non_local(device, cfg0):
    s, psi, osys = cfg0
    kA = (0.1, 0.2, -0.15)
    say("projector", kA)
    ham = inq_amd.ks_hamiltonian(s, max_bands=16)
    ham.uniform_vector_potential = kA
    d = to_dev(psi, device)
    sph = ham.projectors_all().project(d)
    sref = osys.project(psi, kA)
    assert relerr(to_host(sph), sref) < TOL
    base = np.random.default_rng(0).standard_normal(psi.shape) + 0j
    acc = ham.projectors_all().apply(sph, to_dev(base, device))
    assert relerr(to_host(acc), osys.apply(sref, base, kA)) < TOL
    vnl = to_host(ham.non_local(d))
    assert relerr(vnl, osys.apply(sref, np.zeros_like(psi), kA)) < TOL
    ham.close()


def test_hpsi_locally_empty_projectors(device, cfg0):
    """projector::locally_empty (projector.hpp:89-91): sphere 2 with no points
    (its table rows stay padding, -1) and sphere 5 with nlm = 0.  Hpsi (k+A
    != 0) against the oracle on the same tables (1e-12), and against the
    library's Hpsi of the system without those two spheres (1e-14); the
    oracle's result for the emptied tables equals its result without the two
    spheres exactly, and differs from the full system's by 1.5e-4."""
    import copy
    s, psi, _ = cfg0
    kA = (0.05, -0.02, 0.03)
    say("locally empty projectors")
    t = dict(s.tables)
    t["sphere_size"] = t["sphere_size"].copy()
    t["sphere_size"][2] = 0
    t["nlm"] = t["nlm"].copy()
    t["nlm"][5] = 0
    se = copy.copy(s)
    se.tables = t
    keep = [i for i in range(t["nproj"]) if i not in (2, 5)]
    td = {k: (v[keep] if hasattr(v, "shape") and v.shape[:1] == (t["nproj"],) else v)
          for k, v in s.tables.items()}
    td["nproj"] = len(keep)
    sd = copy.copy(s)
    sd.tables = td
    out = []
    for sys_ in (se, sd):
        ham = inq_amd.ks_hamiltonian(sys_, max_bands=16)
        ham.uniform_vector_potential = kA
        out.append(to_host(ham(to_dev(psi, device))))
        ham.close()
    ref = O.System(se.n, se.lattice, se.vloc, se.tables).hamiltonian(psi, kA)
    assert relerr(out[0], ref) < TOL
    assert band_relerr(out[0], ref) < TOL
    assert relerr(out[0], out[1]) < 1e-14


def test_hpsi_bitwise_deterministic(device, cfg0):
    s, psi, _ = cfg0
    say("determinism")
    ham = inq_amd.ks_hamiltonian(s, max_bands=16)
    d = to_dev(psi, device)
    a = to_host(ham(d))
    b = to_host(ham(d))
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    ham.close()


def test_harmonic_oscillator_known_answer_gpu(device):
    """ks_hamiltonian.hpp:356-397 on the GPU: mean |H phi - 1.5 w phi| = 0.0051420503."""
    say("harmonic oscillator")
    lat = np.eye(3) * 10.0
    n = (20, 20, 20)
    r = grid_positions(lat, n)
    r2 = (r * r).sum(-1)
    s = systems.System(n=n, lattice=lat, vloc=2.0 * r2, atoms=np.zeros((0, 3)),
                       tables={"nproj": 0})
    phi = np.repeat(np.exp(-2.0 * r2)[..., None], 6, axis=-1).astype(complex)
    ham = inq_amd.ks_hamiltonian(s, max_bands=6)
    h = to_host(ham(to_dev(phi, device)))
    assert np.abs(h - 3.0 * phi).mean() == pytest.approx(0.0051420503, abs=1e-10)
    ham.close()


@pytest.mark.parametrize("order", [1, 4])
def test_exponential_in_place_matches_oracle(device, cfg0, order):
    s, psi, osys = cfg0
    kA = (0.0, 0.0, 0.07)
    say("exponential order", order)
    ham = inq_amd.ks_hamiltonian(s, max_bands=16)
    ham.uniform_vector_potential = kA
    d = to_dev(psi[..., :6], device)
    operations.exponential_in_place(ham, 0.025j, d, order)
    ref = osys.exponential_in_place(psi[..., :6], 0.025j, order, kA)
    assert relerr(to_host(d), ref) < TOL
    ham.close()


def test_exponential_2_for_1_matches_oracle(device, cfg0):
    s, psi, osys = cfg0
    say("exponential_2_for_1")
    ham = inq_amd.ks_hamiltonian(s, max_bands=16)
    d = to_dev(psi[..., :4], device)
    e = operations.exponential_2_for_1(ham, 0.025j, 0.05j, d, 4)
    rphi, rexp = osys.exponential_2_for_1(psi[..., :4], 0.025j, 0.05j, 4)
    assert relerr(to_host(d), rphi) < TOL
    assert relerr(to_host(e), rexp) < TOL
    ham.close()


def test_etrs_steps_match_oracle(device, cfg0):
    s, psi, osys = cfg0
    say("etrs")
    lat = s.lattice
    pert = inq_amd.real_time.fourier_series_vector_potential([0.03, -0.02, 0.01], 0.1)
    ham = inq_amd.ks_hamiltonian(s, max_bands=16)
    d = to_dev(psi[..., :4], device)
    ref = psi[..., :4].copy()
    dt = 0.05
    for step in range(3):
        t = step * dt
        inq_amd.real_time.etrs(ham, d, t, dt, pert)
        At = lat @ pert.uniform_vector_potential(t)
        Atdt = lat @ pert.uniform_vector_potential(t + dt)
        ref = osys.etrs_step(ref, dt, At, Atdt)
    err = relerr(to_host(d), ref)
    say("etrs rel err after 3 steps", err)
    assert err < TOL
    ham.close()


@pytest.fixture(scope="module")
def cfg1():
    s = systems.diamond_system(2)
    psi = systems.random_orbitals(s, 256)
    return s, psi


def test_hpsi_config1_full_size(device, cfg1):
    """64^3 x 256 bands, 64 atoms: oracle parity on a band sample (bands are
    independent under H), plus size-independent properties on all bands:
    linearity and Hermiticity."""
    s, psi = cfg1
    say("hpsi config1 full size")
    ham = inq_amd.ks_hamiltonian(s, max_bands=256)
    d = to_dev(psi, device)
    h = ham(d)
    sample = [0, 1, 127, 255]
    osys = O.System(s.n, s.lattice, s.vloc, s.tables)
    ref = osys.hamiltonian(np.ascontiguousarray(psi[..., sample]))
    got = to_host(h)[..., sample]
    err = relerr(got, ref)
    say("config1 sample rel err", err)
    assert err < TOL and band_relerr(got, ref) < TOL
    # linearity: H(a psi + b chi) = a H psi + b H chi, with chi = psi rolled by one band
    a, b = 0.7 - 0.2j, -1.3 + 0.4j
    chi = torch.roll(d, 1, dims=-1)
    lhs = ham(a * d + b * chi)
    rhs = a * h + b * torch.roll(h, 1, dims=-1)
    assert (torch.linalg.vector_norm(lhs - rhs) / torch.linalg.vector_norm(rhs)).item() < TOL
    # Hermiticity: <psi_i|H psi_j> = conj(<psi_j|H psi_i>) over all bands
    m = (d.reshape(-1, 256).conj().T.contiguous() @ h.reshape(-1, 256)) * (s.volume / s.npoints)
    herm = (torch.linalg.matrix_norm(m - m.conj().T) / torch.linalg.matrix_norm(m)).item()
    say("config1 hermiticity", herm)
    assert herm < TOL
    ham.close()


def test_etrs_graph_replays_match_eager_bitwise(device, cfg0):
    """hipGraph-captured ETRS step: replayed 4 times with eager work and host
    synchronisations between the replays; every replay equals the eager step."""
    s, psi, osys = cfg0
    say("etrs graph vs eager")
    pert = inq_amd.real_time.fourier_series_vector_potential([0.03, -0.02, 0.01], 0.1)
    dt = 0.05
    ham_g = inq_amd.ks_hamiltonian(s, max_bands=16)
    ham_e = inq_amd.ks_hamiltonian(s, max_bands=16)
    dg = to_dev(psi[..., :8], device)
    de = dg.clone()
    g = inq_amd.real_time.etrs_graph(ham_g, dg, dt, 4, pert=None)
    g.pert = pert  # host-supplied A(t) mode
    for step in range(4):
        t = step * dt
        g.launch(t)
        inq_amd.real_time.etrs(ham_e, de, t, dt, pert)
        junk = ham_e(de)  # unrelated eager work between replays
        torch.cuda.synchronize()
        a, b = to_host(dg), to_host(de)
        say("replay", step, "max |diff|", float(np.abs(a - b).max()))
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
        del junk
    g.close()
    ham_g.close()
    ham_e.close()


@pytest.mark.parametrize("nst", [4, 64])
def test_hpsi_48_matches_oracle(device, nst):
    """The a0 cell on a 48^3 grid (8 atoms, 2501-point spheres): few
    (sphere, band tile) pairs, so the projector splits each sphere over
    point chunks."""
    s = systems.diamond_system(1, grid=48)
    psi = systems.random_orbitals(s, nst)
    say("hpsi 48^3 x", nst)
    ham = inq_amd.ks_hamiltonian(s, max_bands=nst)
    ham.uniform_vector_potential = (0.02, -0.01, 0.03)
    d = to_dev(psi, device)
    h = to_host(ham(d))
    k = min(nst, 4)
    ref = O.System(s.n, s.lattice, s.vloc, s.tables).hamiltonian(psi[..., :k], (0.02, -0.01, 0.03))
    err = relerr(h[..., :k], ref)
    vnl = to_host(ham.non_local(d))[..., :k]
    ref_nl = O.System(s.n, s.lattice, np.zeros(s.n), s.tables).hamiltonian(psi[..., :k], (0.02, -0.01, 0.03)) \
        - O.System(s.n, s.lattice, np.zeros(s.n)).hamiltonian(psi[..., :k], (0.02, -0.01, 0.03))
    say("hpsi rel err %.2e, V_nl rel err %.2e" % (err, relerr(vnl, ref_nl)))
    assert err < TOL
    assert relerr(vnl, ref_nl) < 1e-11
    ham.close()


def test_fourier_space_hpsi_matches_oracle(device, cfg0):
    """ks_hamiltonian::operator() on G-space orbitals against the oracle's
    restatement in the reference's order (ks_hamiltonian.hpp:156-176)."""
    s, psi, osys = cfg0
    say("fourier-space hpsi")
    phi_g = O.to_fourier(psi[..., :5])
    ham = inq_amd.ks_hamiltonian(s, max_bands=16)
    kA = (0.04, -0.02, 0.03)
    ham.uniform_vector_potential = kA
    h = to_host(ham(to_dev(phi_g, device), fourier=True))
    assert relerr(h, osys.hamiltonian_fourier(phi_g, kA)) < TOL
    ham.close()


@pytest.mark.parametrize("L", [48, 64])
def test_fourier_space_hpsi_sweeps_match_oracle(device, L):
    """The G-space Hpsi's three-sweep form (to_real, projector, then
    to_fourier of V psi + V_nl psi with 1/2 |G+k+A|^2 phi added in the last
    sweep) on a radix-6 and a radix-8 grid, with a vector potential, two
    band chunks at 64, against the oracle's restatement."""
    s = systems.diamond_system(1, grid=48) if L == 48 else systems.diamond_system(2)
    nst = 24 if L == 48 else 72
    psi = systems.random_orbitals(s, nst, seed=2)
    phi_g = O.to_fourier(psi)
    kA = (0.03, -0.02, 0.05)
    say("fourier-space hpsi sweeps %d^3 x %d" % (L, nst))
    ham = inq_amd.ks_hamiltonian(s, max_bands=nst)
    ham.uniform_vector_potential = kA
    h = to_host(ham(to_dev(phi_g, device), fourier=True))
    k = 5
    ref = O.System(s.n, s.lattice, s.vloc, s.tables).hamiltonian_fourier(
        np.ascontiguousarray(phi_g[..., -k:]), kA)
    assert relerr(h[..., -k:], ref) < TOL
    ham.close()


def test_energy_terms_match_oracle(device, cfg0):
    """energy::calculate's per-band terms (energy.hpp:59-87): eigenvalues,
    residual norms and non-local expectation values against the oracle, with
    a vector potential, and the occupation-weighted sums."""
    s, psi, osys = cfg0
    say("energy terms")
    nst = 7
    phi = psi[..., :nst]
    ham = inq_amd.ks_hamiltonian(s, max_bands=16)
    kA = (0.03, 0.01, -0.02)
    ham.uniform_vector_potential = kA
    eps, nr, nl = (to_host(t) for t in ham.energy(to_dev(phi, device)))
    reps, rnr, rnl = osys.energy(phi, kA)
    say("energy max abs diffs %.1e %.1e %.1e" % (np.abs(eps - reps).max(),
                                                 np.abs(nr - rnr).max(), np.abs(nl - rnl).max()))
    assert relerr(eps, reps) < TOL
    assert relerr(nr, rnr) < TOL
    assert relerr(nl, rnl) < TOL
    occ = np.linspace(2.0, 0.5, nst)
    e_sum, nl_sum, _ = ham.energy(to_dev(phi, device), occ=occ)
    assert abs(e_sum - occ @ reps) < TOL * abs(occ @ reps)
    assert abs(nl_sum - occ @ rnl) < TOL * abs(occ @ rnl)
    ham.close()


def test_hpsi_96_matches_oracle(device):
    """The 3x3x3 supercell grid (96^3, 216 atoms): three-sweep Hpsi with the
    1024-thread z forward tiles and the 32-column z_bwd_epi tiles (one band
    tile, 3 of its 8 bands valid), 3 bands against the oracle."""
    s = systems.diamond_system(3)
    psi = systems.random_orbitals(s, 3)
    say("hpsi 96^3 x 3")
    ham = inq_amd.ks_hamiltonian(s, max_bands=3)
    ham.uniform_vector_potential = (0.02, -0.01, 0.03)
    h = to_host(ham(to_dev(psi, device)))
    ref = O.System(s.n, s.lattice, s.vloc, s.tables).hamiltonian(psi, (0.02, -0.01, 0.03))
    assert relerr(h, ref) < TOL
    assert band_relerr(h, ref) < TOL
    ham.close()


@pytest.mark.parametrize("L", [60, 72])
def test_hpsi_60_72_match_oracle(device, L):
    """Grids whose z tiles and x-y planes run radix-12 (60 = 12 x 5) and
    radix-9 (72 = 8 x 9) butterflies, the 72-long tiles with 768 threads:
    Hpsi of 11 bands with a vector potential, the batch kernels on the same
    system, and a Taylor-4 exponential (the accumulating epilogue) against
    the oracle."""
    from inq_amd import batch
    s = systems.uneven_batch_system(7, L, 6)
    psi = systems.random_orbitals(s, 11, seed=3, system_id=L)
    kA = (0.02, -0.01, 0.03)
    say("hpsi %d^3 x 11" % L)
    ham = inq_amd.ks_hamiltonian(s, max_bands=11)
    ham.uniform_vector_potential = kA
    d = to_dev(psi, device)
    h = to_host(ham(d))
    osys = O.System(s.n, s.lattice, s.vloc, s.tables)
    ref = osys.hamiltonian(psi, kA)
    assert relerr(h, ref) < TOL
    assert band_relerr(h, ref) < TOL
    b = batch.hamiltonian_batch([ham])
    for group in (False, True):
        say("batch kernels %d^3, one launch per grid length: %s" % (L, group))
        b.set_tuning(group_by_grid=group)
        out = b([d])[0]
        torch.cuda.synchronize()
        assert relerr(to_host(out), ref) < TOL
    b.close()
    say("exponential %d^3 x 3" % L)
    e = to_dev(psi[..., :3], device)
    operations.exponential_in_place(ham, 0.025j, e, 4)
    assert relerr(to_host(e), osys.exponential_in_place(psi[..., :3], 0.025j, 4, kA)) < TOL
    ham.close()


@pytest.mark.parametrize("L", [24, 30, 36, 40, 80])
def test_hpsi_remaining_grids_match_oracle(device, L):
    """The other grid lengths with three-sweep kernels (24 = 8 x 3, 30 = 6 x
    5, 36 = 6 x 6, 40 = 8 x 5, and 80 = 8 x 2 x 5 with 1024-thread z tiles):
    Hpsi of 5 bands (one ragged band tile) with a vector potential, and a
    Taylor-4 exponential of 3 bands, against the oracle."""
    s = systems.uneven_batch_system(11, L, 4)
    psi = systems.random_orbitals(s, 5, seed=5, system_id=L)
    kA = (-0.03, 0.02, 0.01)
    say("hpsi %d^3 x 5" % L)
    ham = inq_amd.ks_hamiltonian(s, max_bands=5)
    ham.uniform_vector_potential = kA
    h = to_host(ham(to_dev(psi, device)))
    osys = O.System(s.n, s.lattice, s.vloc, s.tables)
    ref = osys.hamiltonian(psi, kA)
    assert relerr(h, ref) < TOL
    assert band_relerr(h, ref) < TOL
    say("exponential %d^3 x 3" % L)
    e = to_dev(psi[..., :3], device)
    operations.exponential_in_place(ham, 0.025j, e, 4)
    assert relerr(to_host(e), osys.exponential_in_place(psi[..., :3], 0.025j, 4, kA)) < TOL
    ham.close()


def test_band_group_chunks_match_contiguous_bitwise(device, cfg1):
    """64^3 x 256 at ld = 256 runs its chunks as interleaved 8-band groups;
    the same bands at ld = 264 run contiguous chunks.  Hpsi, G-space Hpsi,
    an ETRS step (Taylor accumulators) and a captured ETRS step must agree
    bitwise: every band goes through the same arithmetic whichever chunk
    holds it."""
    s, psi = cfg1
    say("band-group chunks vs contiguous, 64^3 x 256")
    ham = inq_amd.ks_hamiltonian(s, max_bands=256)
    nst, ldp = 256, 264
    d = to_dev(psi, device)
    pad = torch.zeros(psi.shape[:3] + (ldp,), dtype=torch.complex128, device=device)
    pad[..., :nst] = d
    hp = torch.zeros_like(pad)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())
    h = ham(d)
    ham.handle.call("inqamd_hamiltonian_apply", ptr(pad), ptr(hp), nst, ldp)
    torch.cuda.synchronize()
    assert torch.equal(h, hp[..., :nst])
    assert torch.count_nonzero(hp[..., nst:]).item() == 0
    A0 = (ctypes.c_double * 3)(0.01, -0.02, 0.03)
    A1 = (ctypes.c_double * 3)(0.012, -0.018, 0.031)
    for t, ld in ((d, nst), (pad, ldp)):
        ham.handle.call("inqamd_etrs_step", ctypes.c_double(0.02), A0, A1, 4, ptr(t), nst, ld)
    torch.cuda.synchronize()
    assert torch.equal(d, pad[..., :nst])
    # G-space Hpsi (its sweeps take the same chunks)
    ham.handle.call("inqamd_hamiltonian_apply_fourier", ptr(d), ptr(h), nst, nst)
    ham.handle.call("inqamd_hamiltonian_apply_fourier", ptr(pad), ptr(hp), nst, ldp)
    torch.cuda.synchronize()
    assert torch.equal(h, hp[..., :nst])
    # a captured step (grouped chunks, A = 0) against the eager contiguous one
    zero = (ctypes.c_double * 3)(0.0, 0.0, 0.0)
    g = inq_amd.real_time.etrs_graph(ham, d, 0.02, 4, pert=None)
    for _ in range(2):
        g.launch(0.0)
        ham.handle.call("inqamd_etrs_step", ctypes.c_double(0.02), zero, zero, 4, ptr(pad), nst, ldp)
        torch.cuda.synchronize()
        assert torch.equal(d, pad[..., :nst])
    g.close()
    ham.close()


def test_band_groups_uneven_chunks_bitwise(device, cfg1):
    """232 bands in rows of 256 (ld = 256): 29 groups of 8 over five chunks
    (6, 6, 6, 6, 5 groups), against the same bands at ld = 264 (contiguous
    chunks): Hpsi bitwise, and the padding columns untouched."""
    s, psi = cfg1
    say("band-group chunks, 232 of 256 columns")
    nst = 232
    ham = inq_amd.ks_hamiltonian(s, max_bands=256)
    src = to_dev(psi[..., :nst], device)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())
    outs = []
    for ld in (256, 264):
        a = torch.zeros(psi.shape[:3] + (ld,), dtype=torch.complex128, device=device)
        a[..., :nst] = src
        o = torch.zeros_like(a)
        ham.handle.call("inqamd_hamiltonian_apply", ptr(a), ptr(o), nst, ld)
        torch.cuda.synchronize()
        assert torch.count_nonzero(o[..., nst:]).item() == 0
        outs.append(o[..., :nst])
    assert torch.equal(outs[0], outs[1])
    ref = O.System(s.n, s.lattice, s.vloc, s.tables).hamiltonian(np.ascontiguousarray(psi[..., [0, 231]]))
    got = to_host(outs[0])[..., [0, 231]]
    assert relerr(got, ref) < TOL
    ham.close()


def test_pipelined_chunks_graph_and_oracle(device):
    """A 64^3 system whose Hpsi runs two band chunks on two streams (max_bands
    64, 40 bands, 32-band chunks): Hpsi against the oracle, and a captured
    ETRS step replayed twice against the eager steps, bitwise."""
    s = systems.diamond_system(2)
    psi = systems.random_orbitals(s, 40)
    say("pipelined chunks 64^3 x 40")
    ham = inq_amd.ks_hamiltonian(s, max_bands=64)
    d = to_dev(psi, device)
    h = to_host(ham(d))
    ref = O.System(s.n, s.lattice, s.vloc, s.tables).hamiltonian(psi[..., :3])
    assert relerr(h[..., :3], ref) < TOL
    pert = inq_amd.real_time.fourier_series_vector_potential([0.03, -0.02], 0.1)
    ham_e = inq_amd.ks_hamiltonian(s, max_bands=64)
    dg, de = d.clone(), d.clone()
    g = inq_amd.real_time.etrs_graph(ham, dg, 0.02, 4, pert=None)
    g.pert = pert
    for step in range(2):
        g.launch(step * 0.02)
        inq_amd.real_time.etrs(ham_e, de, step * 0.02, 0.02, pert)
        torch.cuda.synchronize()
        assert np.array_equal(to_host(dg).view(np.uint64), to_host(de).view(np.uint64))
    g.close()
    ham.close()
    ham_e.close()


def test_etrs_series_graph_matches_oracle(device, cfg0):
    """Series-mode graph: A(t) evaluated on the device from a step counter.
    dt = 0.02 keeps Taylor-4 inside its stability bound on this grid
    (|lambda_max| dt / 2 = 1.44 < 2 sqrt 2), so 8 steps stay comparable."""
    s, psi, osys = cfg0
    say("etrs series graph")
    pert = inq_amd.real_time.fourier_series_vector_potential([0.04, -0.03, 0.02, 0.01], 0.1)
    dt = 0.02
    ham = inq_amd.ks_hamiltonian(s, max_bands=16)
    d = to_dev(psi[..., :4], device)
    g = inq_amd.real_time.etrs_graph(ham, d, dt, 4, pert=pert)
    g.set_step(0)
    ref = psi[..., :4].copy()
    lat = s.lattice
    for step in range(8):
        g.launch()
        t = step * dt
        ref = osys.etrs_step(ref, dt, lat @ pert.uniform_vector_potential(t),
                             lat @ pert.uniform_vector_potential(t + dt))
        err = relerr(to_host(d), ref)
        say("series step", step, "rel err", err)
        assert err < TOL
    g.close()
    ham.close()


def test_aa_mfma_f64_lane_layout(device):
    """The FP64 matrix-core lane layout the projector kernels assume."""
    say("mfma selftest")
    out = torch.zeros(256, dtype=torch.float64, device=device)
    L = inq_amd._lib.lib()
    inq_amd._lib.check(L.inqamd_selftest_mfma(ctypes.c_void_p(out.data_ptr()),
                                              ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    got = out.cpu().numpy().reshape(16, 16)
    i = np.arange(16)[:, None]
    k = np.arange(4)[None, :]
    A = 1.0 + i + 16.0 * k
    B = 2.0 + 3.0 * np.arange(16)[None, :] - 5.0 * np.arange(4)[:, None]
    assert np.array_equal(got, A @ B)


def test_aa_mfma_f64_4x4x4_lane_layout(device):
    """The lane layout of the 4-block 4x4x4 FP64 matrix instruction the
    projector coefficients of the three-sweep Hpsi use."""
    say("mfma 4x4x4 selftest")
    out = torch.zeros(64, dtype=torch.float64, device=device)
    L = inq_amd._lib.lib()
    inq_amd._lib.check(L.inqamd_selftest_mfma_4x4(
        ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    got = out.cpu().numpy().reshape(4, 4, 4)
    i = np.arange(4)[:, None]
    k = np.arange(4)[None, :]
    for s in range(4):
        A = 1.0 + i + 4.0 * k + 16.0 * s
        B = 2.0 + 3.0 * np.arange(4)[None, :] - 5.0 * np.arange(4)[:, None] + 7.0 * s
        assert np.array_equal(got[s], A @ B), s


@pytest.mark.parametrize("env", [{"INQAMD_STREAMS": "3"}, {"INQAMD_BAND_GROUP_SIZE": "16"},
                                 {"INQAMD_BAND_GROUP_SIZE": "32", "INQAMD_BAND_CHUNK": "64"}],
                         ids=["three-streams", "groups-16", "groups-32-chunk-64"])
def test_pipeline_variants_in_fresh_process(device, env):
    """The library reads these switches once per process, so each variant runs
    in a child process: grouped chunks (ld = 256) against contiguous ones
    (ld = 264) bitwise for Hpsi, an eager and a captured ETRS step, plus Hpsi
    against the oracle (tests/env_variant_check.py)."""
    import os
    import subprocess
    say("pipeline variant", env)
    e = dict(os.environ)
    e.update(env)
    script = os.path.join(os.path.dirname(os.path.abspath(__file__)), "env_variant_check.py")
    r = subprocess.run([sys.executable, "-u", script], env=e, capture_output=True, text=True,
                       timeout=240)
    say(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and "ENV-VARIANT-OK" in r.stdout


def test_hpsi_96_z_tile_shapes(device):
    """96^3, 20 bands (two whole 8-band tiles and a ragged one): the 32-column
    z_bwd_epi tiles of 4 y x 8 bands (default) and 8 y x 4 bands (z_tile_y =
    16) give the same bits, and bands 0 and 19 match the oracle."""
    s = systems.diamond_system(3)
    nst = 20
    say("hpsi 96^3 x 20, z tile shapes")
    g = torch.Generator(device=device).manual_seed(960)
    d = torch.randn(tuple(s.n) + (nst,), dtype=torch.complex128, device=device, generator=g)
    ham = inq_amd.ks_hamiltonian(s, max_bands=nst, tuning=None)
    ham.uniform_vector_potential = (0.01, 0.02, -0.03)
    h8 = ham(d)
    ham.tuning = {"z_tile_y": 16}
    h16 = ham(d)
    torch.cuda.synchronize()
    assert torch.equal(h8, h16)
    x = np.ascontiguousarray(d[..., [0, 19]].cpu().numpy())
    ref = O.System(s.n, s.lattice, s.vloc, s.tables).hamiltonian(x, (0.01, 0.02, -0.03))
    err = relerr(h8[..., [0, 19]].cpu().numpy(), ref)
    say("96^3 Hpsi vs oracle on 2 of 20 bands %.2e" % err)
    assert err < TOL
    ham.close()


def test_band_groups_96_match_padded_rows(device):
    """96^3 x 512 (config 4's shape, 216 atoms): the default chunks of 24
    bands take interleaved 8-band groups at ld = 512 and contiguous ranges at
    ld = 520; Hpsi must agree bitwise, and two bands against the oracle."""
    s = systems.diamond_system(3)
    nst, ldp = 512, 520
    say("band groups 96^3 x 512 vs ld 520")
    g = torch.Generator(device=device).manual_seed(96)
    d = torch.randn(tuple(s.n) + (nst,), dtype=torch.complex128, device=device, generator=g)
    ham = inq_amd.ks_hamiltonian(s, max_bands=nst)
    h = ham(d)
    pad = torch.zeros(tuple(s.n) + (ldp,), dtype=torch.complex128, device=device)
    pad[..., :nst] = d
    del d
    hp = torch.zeros_like(pad)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())
    ham.handle.call("inqamd_hamiltonian_apply", ptr(pad), ptr(hp), nst, ldp)
    torch.cuda.synchronize()
    assert torch.equal(h, hp[..., :nst])
    assert torch.count_nonzero(hp[..., nst:]).item() == 0
    two = h[..., [3, 509]].cpu().numpy()
    x = np.ascontiguousarray(pad[..., [3, 509]].cpu().numpy())
    ref = O.System(s.n, s.lattice, s.vloc, s.tables).hamiltonian(x)
    err = relerr(two, ref)
    say("96^3 Hpsi vs oracle on 2 bands %.2e" % err)
    assert err < TOL
    ham.close()


def test_tuning_knobs_do_not_change_results(device):
    """The z-tile shape, band chunk and projector bands of inqamd_tuning leave
    Hpsi and an ETRS step the same bit for bit; the x-y thread count is
    another instantiation of the plane kernel, whose floating-point
    contraction may differ (1e-14 relative); invalid values are refused."""
    s = systems.diamond_system(1, grid=48)
    ham = inq_amd.ks_hamiltonian(s, max_bands=64, tuning=None)
    ham.uniform_vector_potential = (0.02, -0.01, 0.03)
    g = torch.Generator(device=device).manual_seed(21)
    psi = torch.randn(tuple(s.n) + (64,), dtype=torch.complex128, device=device, generator=g)
    ref = ham(psi)
    step = psi.clone()
    ham.handle.call("inqamd_etrs_step", ctypes.c_double(0.01), inq_amd._lib.dvec([0.01, 0, 0]),
                    inq_amd._lib.dvec([0.012, 0, 0]), 4, ctypes.c_void_p(step.data_ptr()), 64, 64)
    torch.cuda.synchronize()
    for knobs in ({"z_tile_y": 16}, {"band_chunk": 16}, {"projector_bands": 64},
                  {"projector_bands": 16, "band_chunk": 24, "z_tile_y": 16}, {"xy_threads": 1}):
        say("tuning", knobs)
        ham.tuning = dict({"z_tile_y": 0, "xy_threads": 0, "band_chunk": 0, "projector_bands": 0},
                          **knobs)
        ham.uniform_vector_potential = (0.02, -0.01, 0.03)  # the ETRS step leaves A(t + dt)
        got = ham(psi)
        other = psi.clone()
        ham.handle.call("inqamd_etrs_step", ctypes.c_double(0.01), inq_amd._lib.dvec([0.01, 0, 0]),
                        inq_amd._lib.dvec([0.012, 0, 0]), 4, ctypes.c_void_p(other.data_ptr()),
                        64, 64)
        torch.cuda.synchronize()
        e1 = float((torch.linalg.vector_norm(got - ref) / torch.linalg.vector_norm(ref)).item())
        e2 = float((torch.linalg.vector_norm(other - step) / torch.linalg.vector_norm(step)).item())
        say("  Hpsi rel diff %.2e, ETRS rel diff %.2e" % (e1, e2))
        if "xy_threads" in knobs:
            assert e1 < 1e-14 and e2 < 1e-14, knobs
        else:
            assert torch.equal(got, ref) and torch.equal(other, step), knobs
    for bad in ({"z_tile_y": 12}, {"band_chunk": 20}, {"band_chunk": 72}, {"projector_bands": 8}):
        with pytest.raises(inq_amd._lib.InqAmdError):
            ham.tuning = bad
    ham.close()
