"""GPU parity of the current observable and of the propagation adjoint.

Current: the HIP library against the oracle (1e-12 relative to |J|) and
against the reference's own known answers (current.hpp:80-177, 1e-9 abs).
Propagation: the forward loss against the oracle's td_current (1e-11
relative); the adjoint gradients (coefficients, V at sampled points, and the
adjoint state at t = 0 along a random direction) against central finite
differences of the ORACLE's loss with h = 1e-5, within 1e-6 relative to the
largest gradient component.
"""
import json
import os
import sys

import numpy as np
import pytest
import torch

import oracle as O
from inq_amd import adjoint, hamiltonian, systems

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
KNOWN = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def to_dev(a, device):
    return torch.from_numpy(np.ascontiguousarray(a, np.complex128)).to(device)


@pytest.fixture(scope="module")
def cfg0():
    s = systems.diamond_system(1)
    psi = systems.random_orbitals(s, 16)
    return s, psi, O.System(s.n, s.lattice, s.vloc, s.tables)


@pytest.mark.parametrize("kA", [(0.0, 0.0, 0.0), (0.05, -0.02, 0.03)])
def test_current_matches_oracle(device, cfg0, kA):
    say("current config0", kA)
    s, psi, osys = cfg0
    ham = hamiltonian.ks_hamiltonian(s, 16)
    ham.uniform_vector_potential = kA
    jc, jx = ham.current(to_dev(psi, device), occ=2.0)
    rc, rx = osys.current(psi, kA, 2.0)
    assert np.abs(jx - rx).max() <= 1e-12 * np.abs(rx).max()
    assert np.abs(jc - rc).max() <= 1e-12 * np.abs(rc).max()
    ham.close()


@pytest.mark.parametrize("case", ["gamma", "gamma_A", "k_quarter", "k_quarter_A"])
def test_current_reference_known_answers_gpu(device, case):
    k = KNOWN["current"]
    say("current known answer", case)
    cell = k["cell_orthorhombic"]
    n = tuple(k["grid"])
    nst = k["bands"]
    phi = O.randomize(n, nst, 0)
    C = phi.reshape(-1, nst)
    dV = np.prod(cell) / np.prod(n)
    L = np.linalg.cholesky(dV * (C.conj().T @ C))
    phi = (C @ np.linalg.inv(L).conj().T).reshape(n + (nst,))
    sysm = systems.System(n=n, lattice=np.diag(cell), vloc=np.zeros(n), atoms=np.zeros((0, 3)),
                          tables={"nproj": 0})
    ham = hamiltonian.ks_hamiltonian(sysm, nst)
    c = k[case]
    ham.uniform_vector_potential = c["A_cov"]
    if "k_frac" in c:
        ham.kpoint = 2.0 * np.pi * np.array(c["k_frac"])
    jc, _ = ham.current(to_dev(phi, device), occ=k["occupation"])
    assert np.allclose(jc, c["J_cov"], rtol=0, atol=1e-9)
    ham.close()


def _small_case(kind):
    """'pencil': a non-cubic grid (five pencil passes per Hpsi); 'cubic': a
    24^3 grid (the three-sweep Hpsi)."""
    if kind == "pencil":
        lat, n = np.diag([6.0, 6.3, 5.7]), (16, 16, 15)
    else:
        lat, n = np.eye(3) * 6.5, (24, 24, 24)
    atoms = np.array([[0.3, -0.2, 0.1], [2.7, 1.9, -1.4]])
    s = systems.make_system(n, lat, atoms, seed=4, system_id=21)
    psi = systems.random_orbitals(s, 3, ecut=5.0)
    return s, psi


PROP = dict(nsteps=5, dt=0.02, omega=0.9, direction=(0.6, 0.0, 0.8), order=4, occ=2.0)
COEF = np.array([0.04, -0.03, 0.025])


@pytest.fixture(scope="module", params=["pencil", "cubic"])
def small(request):
    s, psi = _small_case(request.param)
    return s, psi, O.System(s.n, s.lattice, s.vloc, s.tables)


def _oracle_loss(s, psi, coef, vloc=None):
    osys = O.System(s.n, s.lattice, s.vloc if vloc is None else vloc, s.tables)
    loss, _ = osys.td_current(psi, coef, PROP["omega"], PROP["direction"], PROP["dt"],
                              PROP["nsteps"], PROP["order"], PROP["occ"])
    return loss


def _prop(ham, nst, **kw):
    p = dict(checkpoint_every=2, band_slice=0, store_powers=True)
    p.update(kw)
    return adjoint.propagation_adjoint(ham, nst, PROP["nsteps"], PROP["dt"], len(COEF),
                                       PROP["omega"], PROP["direction"], PROP["order"],
                                       PROP["occ"], **p)


def test_td_forward_loss_matches_oracle(device, small):
    say("td forward small")
    s, psi, osys = small
    ham = hamiltonian.ks_hamiltonian(s, 3)
    prop = _prop(ham, 3)
    psiT = torch.empty_like(to_dev(psi, device))
    loss, _ = prop.forward(COEF, to_dev(psi, device), psiT)
    ref, ref_psi = osys.td_current(psi, COEF, PROP["omega"], PROP["direction"], PROP["dt"],
                                   PROP["nsteps"], PROP["order"], PROP["occ"])
    torch.cuda.synchronize()
    err = np.linalg.norm(psiT.cpu().numpy() - ref_psi) / np.linalg.norm(ref_psi)
    assert err < 1e-12
    assert abs(loss - ref) <= 1e-11 * max(abs(ref), 1.0)
    prop.close()
    ham.close()


@pytest.fixture(scope="module")
def small_grads(device, small):
    s, psi, _ = small
    ham = hamiltonian.ks_hamiltonian(s, 3)
    prop = _prop(ham, 3)
    prop.forward(COEF, to_dev(psi, device))
    gc, gV, lam = prop.backward(1.0, want_lambda0=True)
    torch.cuda.synchronize()
    out = (gc.copy(), gV.cpu().numpy().reshape(s.n), lam.cpu().numpy())
    prop.close()
    ham.close()
    return out


def test_adjoint_coefficients_match_oracle_fd(device, small, small_grads):
    say("adjoint dL/dc vs oracle FD")
    s, psi, _ = small
    gc = small_grads[0]
    h = 1e-5
    fd = np.zeros_like(COEF)
    for j in range(COEF.size):
        e = np.zeros_like(COEF)
        e[j] = h
        fd[j] = (_oracle_loss(s, psi, COEF + e) - _oracle_loss(s, psi, COEF - e)) / (2 * h)
    scale = np.abs(fd).max()
    assert scale > 1e-6
    assert np.abs(gc - fd).max() <= 1e-6 * scale, (gc, fd)


def test_adjoint_potential_matches_oracle_fd(device, small, small_grads):
    say("adjoint dL/dV vs oracle FD")
    s, psi, _ = small
    gV = small_grads[1]
    h = 1e-5
    rng = np.random.default_rng(7)
    # the largest-gradient point plus seeded samples
    pts = [np.unravel_index(np.argmax(np.abs(gV)), s.n)]
    pts += [tuple(int(rng.integers(0, n)) for n in s.n) for _ in range(3)]
    scale = np.abs(gV).max()
    for pt in pts:
        vp = s.vloc.copy()
        vm = s.vloc.copy()
        vp[pt] += h
        vm[pt] -= h
        fd = (_oracle_loss(s, psi, COEF, vp) - _oracle_loss(s, psi, COEF, vm)) / (2 * h)
        assert abs(gV[pt] - fd) <= 1e-6 * scale, (pt, gV[pt], fd)


def test_adjoint_initial_state_matches_oracle_fd(device, small, small_grads):
    say("adjoint lambda0 vs oracle FD")
    s, psi, _ = small
    lam = small_grads[2]
    rng = np.random.default_rng(3)
    d = rng.standard_normal(psi.shape) + 1j * rng.standard_normal(psi.shape)
    d *= np.linalg.norm(psi) / np.linalg.norm(d)
    h = 1e-5
    fd = (_oracle_loss(s, psi + h * d, COEF) - _oracle_loss(s, psi - h * d, COEF)) / (2 * h)
    an = float(np.real(np.vdot(lam, d)))
    assert abs(an - fd) <= 1e-6 * max(abs(fd), 1e-3), (an, fd)


@pytest.mark.parametrize("kw", [dict(band_slice=2), dict(store_powers=False),
                                dict(checkpoint_every=5), dict(checkpoint_every=1, band_slice=1)])
def test_adjoint_layouts_agree(device, small, small_grads, kw):
    say("adjoint layout", kw)
    s, psi, _ = small
    ham = hamiltonian.ks_hamiltonian(s, 3)
    prop = _prop(ham, 3, **kw)
    prop.forward(COEF, to_dev(psi, device))
    gc, gV, _ = prop.backward(1.0)
    torch.cuda.synchronize()
    gc0, gV0, _ = small_grads
    assert np.abs(gc - gc0).max() <= 1e-12 * np.abs(gc0).max()
    assert np.abs(gV.cpu().numpy().reshape(s.n) - gV0).max() <= 1e-12 * np.abs(gV0).max()
    prop.close()
    ham.close()


def test_autograd_function(device, small, small_grads):
    say("autograd function")
    s, psi, _ = small
    ham = hamiltonian.ks_hamiltonian(s, 3)
    prop = _prop(ham, 3)
    coef = torch.tensor(COEF, dtype=torch.float64, requires_grad=True)
    vloc = torch.from_numpy(s.vloc.copy()).to(device).requires_grad_(True)
    loss = adjoint.td_current_loss(coef, vloc, prop, to_dev(psi, device))
    (2.0 * loss).backward()
    gc0, gV0, _ = small_grads
    assert np.abs(coef.grad.numpy() - 2 * gc0).max() <= 1e-12 * np.abs(gc0).max()
    assert np.abs(vloc.grad.cpu().numpy() - 2 * gV0).max() <= 1e-12 * np.abs(gV0).max()
    prop.close()
    ham.close()
