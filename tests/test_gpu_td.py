"""GPU parity of the current observable and of the propagation adjoint.

Current: the HIP library against the oracle (1e-12 relative to |J|) in both
forms of the [V_nl, r] term -- "inq", position_commutator as the reference
writes it (projector_all.hpp:362-366), and "corrected" -- on config 0 (8
atoms with projectors, k + A != 0), and against the reference's own known
answers (current.hpp:80-177, 1e-9 abs).
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


@pytest.mark.parametrize("mode", ["inq", "corrected"])
@pytest.mark.parametrize("kA", [(0.0, 0.0, 0.0), (0.05, -0.02, 0.03)])
def test_current_matches_oracle(device, cfg0, kA, mode):
    say("current config0", kA, mode)
    s, psi, osys = cfg0
    ham = hamiltonian.ks_hamiltonian(s, 16)
    ham.uniform_vector_potential = kA
    assert ham.position_commutator == "inq"
    ham.position_commutator = mode
    jc, jx = ham.current(to_dev(psi, device), occ=2.0)
    rc, rx = osys.current(psi, kA, 2.0, commutator=mode)
    say("current", mode, jx, "oracle", rx)
    assert np.abs(jx - rx).max() <= 1e-12 * np.abs(rx).max()
    assert np.abs(jc - rc).max() <= 1e-12 * np.abs(rc).max()
    ham.close()


def test_current_forms_differ_on_config0(device, cfg0):
    """The two commutator forms give different currents when projectors
    overlap psi (so the parity test above tells them apart)."""
    s, psi, osys = cfg0
    kA = (0.05, -0.02, 0.03)
    a = osys.current(psi, kA, 2.0, commutator="inq")[1]
    b = osys.current(psi, kA, 2.0, commutator="corrected")[1]
    assert np.abs(a - b).max() > 1e-8 * np.abs(a).max()


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


class Small:
    def __init__(self, kind, mode):
        self.s, self.psi = _small_case(kind)
        self.mode = mode
        self.osys = O.System(self.s.n, self.s.lattice, self.s.vloc, self.s.tables)

    def __iter__(self):
        return iter((self.s, self.psi, self.osys))


# the loss in both commutator forms: "inq" on both Hpsi paths, "corrected" on one
@pytest.fixture(scope="module", params=[("pencil", "inq"), ("cubic", "inq"),
                                        ("cubic", "corrected")],
                ids=["pencil-inq", "cubic-inq", "cubic-corrected"])
def small(request):
    return Small(*request.param)


def _oracle_loss(s, psi, coef, vloc=None, mode="inq"):
    osys = O.System(s.n, s.lattice, s.vloc if vloc is None else vloc, s.tables)
    loss, _ = osys.td_current(psi, coef, PROP["omega"], PROP["direction"], PROP["dt"],
                              PROP["nsteps"], PROP["order"], PROP["occ"],
                              commutator=mode)
    return loss


def _prop(ham, nst, **kw):
    p = dict(checkpoint_every=2, band_slice=0, store_powers=True)
    p.update(kw)
    return adjoint.propagation_adjoint(ham, nst, PROP["nsteps"], PROP["dt"], len(COEF),
                                       PROP["omega"], PROP["direction"], PROP["order"],
                                       PROP["occ"], **p)


def test_td_forward_loss_matches_oracle(device, small):
    say("td forward small", small.mode)
    s, psi, osys = small
    ham = hamiltonian.ks_hamiltonian(s, 3)
    prop = _prop(ham, 3, commutator=small.mode)
    psiT = torch.empty_like(to_dev(psi, device))
    loss, _ = prop.forward(COEF, to_dev(psi, device), psiT)
    ref, ref_psi = osys.td_current(psi, COEF, PROP["omega"], PROP["direction"], PROP["dt"],
                                   PROP["nsteps"], PROP["order"], PROP["occ"],
                                   commutator=small.mode)
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
    ham.position_commutator = small.mode
    prop = _prop(ham, 3)
    assert prop.commutator == small.mode
    prop.forward(COEF, to_dev(psi, device))
    gc, gV, lam = prop.backward(1.0, want_lambda0=True)
    torch.cuda.synchronize()
    out = (gc.copy(), gV.cpu().numpy().reshape(s.n), lam.cpu().numpy())
    prop.close()
    ham.close()
    return out


def test_adjoint_coefficients_match_oracle_fd(device, small, small_grads):
    say("adjoint dL/dc vs oracle FD", small.mode)
    s, psi, _ = small
    gc = small_grads[0]
    h = 1e-5
    fd = np.zeros_like(COEF)
    for j in range(COEF.size):
        e = np.zeros_like(COEF)
        e[j] = h
        fd[j] = (_oracle_loss(s, psi, COEF + e, mode=small.mode) -
                 _oracle_loss(s, psi, COEF - e, mode=small.mode)) / (2 * h)
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
        fd = (_oracle_loss(s, psi, COEF, vp, mode=small.mode) -
              _oracle_loss(s, psi, COEF, vm, mode=small.mode)) / (2 * h)
        assert abs(gV[pt] - fd) <= 1e-6 * scale, (pt, gV[pt], fd)


def test_adjoint_initial_state_matches_oracle_fd(device, small, small_grads):
    say("adjoint lambda0 vs oracle FD")
    s, psi, _ = small
    lam = small_grads[2]
    rng = np.random.default_rng(3)
    d = rng.standard_normal(psi.shape) + 1j * rng.standard_normal(psi.shape)
    d *= np.linalg.norm(psi) / np.linalg.norm(d)
    h = 1e-5
    fd = (_oracle_loss(s, psi + h * d, COEF, mode=small.mode) -
          _oracle_loss(s, psi - h * d, COEF, mode=small.mode)) / (2 * h)
    an = float(np.real(np.vdot(lam, d)))
    assert abs(an - fd) <= 1e-6 * max(abs(fd), 1e-3), (an, fd)


@pytest.mark.parametrize("kw", [dict(band_slice=2), dict(store_powers=False),
                                dict(checkpoint_every=5), dict(checkpoint_every=1, band_slice=1)])
def test_adjoint_layouts_agree(device, small, small_grads, kw):
    say("adjoint layout", kw)
    s, psi, _ = small
    ham = hamiltonian.ks_hamiltonian(s, 3)
    prop = _prop(ham, 3, commutator=small.mode, **kw)
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
    prop = _prop(ham, 3, commutator=small.mode)
    coef = torch.tensor(COEF, dtype=torch.float64, requires_grad=True)
    vloc = torch.from_numpy(s.vloc.copy()).to(device).requires_grad_(True)
    loss = adjoint.td_current_loss(coef, vloc, prop, to_dev(psi, device))
    (2.0 * loss).backward()
    gc0, gV0, _ = small_grads
    assert np.abs(coef.grad.numpy() - 2 * gc0).max() <= 1e-12 * np.abs(gc0).max()
    assert np.abs(vloc.grad.cpu().numpy() - 2 * gV0).max() <= 1e-12 * np.abs(gV0).max()
    prop.close()
    ham.close()


def test_autograd_second_forward_before_backward_raises(device, small):
    """Two live losses on one propagation object: the first one's backward
    would read the second forward's checkpoints, so it raises."""
    if small.mode != "inq" or small.s.n[0] != 24:
        pytest.skip("one case suffices")
    s, psi, _ = small
    ham = hamiltonian.ks_hamiltonian(s, 3)
    prop = _prop(ham, 3)
    coef = torch.tensor(COEF, dtype=torch.float64, requires_grad=True)
    vloc = torch.from_numpy(s.vloc.copy()).to(device).requires_grad_(True)
    l1 = adjoint.td_current_loss(coef, vloc, prop, to_dev(psi, device))
    l2 = adjoint.td_current_loss(coef, vloc, prop, to_dev(psi, device))
    with pytest.raises(RuntimeError, match="another forward"):
        (l1 + l2).backward()
    # the library itself refuses a backward after V changed
    prop.forward(COEF, to_dev(psi, device))
    ham.set_scalar_potential(vloc.detach() + 0.1)
    with pytest.raises(Exception, match="changed since the forward"):
        prop.backward(1.0)
    prop.close()
    ham.close()


@pytest.mark.parametrize("order", [1, 2, 3, 5, 6, 7, 8])
def test_adjoint_taylor_orders(device, order):
    """The fused half-step kernels (dL/dV, the G-space pair sums) take the
    Taylor order as a template argument: at each order from 1 to 8 other
    than the 4 of the tests above, the 24^3 case's dL/dc matches central
    differences of the oracle's loss (h = 1e-5, 1e-6 of the largest), and
    dL/dc and dL/dV match the unfused form (store_powers=False, which keeps
    no powers and runs none of those kernels) to 1e-12."""
    say("adjoint Taylor order", order)
    s, psi = _small_case("cubic")
    ham = hamiltonian.ks_hamiltonian(s, 3)
    out = []
    for store in (True, False):
        prop = adjoint.propagation_adjoint(
            ham, 3, PROP["nsteps"], PROP["dt"], len(COEF), PROP["omega"], PROP["direction"],
            order, PROP["occ"], checkpoint_every=2, band_slice=0, store_powers=store,
            commutator="inq")
        prop.forward(COEF, to_dev(psi, device))
        gc, gV, _ = prop.backward(1.0)
        torch.cuda.synchronize()
        out.append((gc.copy(), gV.cpu().numpy()))
        prop.close()
    ham.close()
    (gc, gV), (gc_u, gV_u) = out
    assert np.abs(gc - gc_u).max() <= 1e-12 * np.abs(gc_u).max()
    assert np.abs(gV - gV_u).max() <= 1e-12 * np.abs(gV_u).max()
    osys = O.System(s.n, s.lattice, s.vloc, s.tables)

    def loss(c):
        v, _ = osys.td_current(psi, c, PROP["omega"], PROP["direction"], PROP["dt"],
                               PROP["nsteps"], order, PROP["occ"], commutator="inq")
        return v

    h = 1e-5
    fd = np.array([(loss(COEF + h * e) - loss(COEF - h * e)) / (2 * h)
                   for e in np.eye(COEF.size)])
    scale = np.abs(fd).max()
    assert scale > 1e-6
    assert np.abs(gc - fd).max() <= 1e-6 * scale, (order, gc, fd)
