"""The differentiation and large-step workloads at their full sizes.

Config 3 (adjoint_48): 48^3 grid, 64 bands, 8 atoms, 32 Fourier coefficients
of A_x(t), Taylor-4 ETRS, 400 steps of dt = 0.015 (the Taylor-4 stability
bound at h = 0.21375 is 0.0175), checkpoints every 20, loss J_x(T) with
INQ's position_commutator.  dL/dc for all 32 coefficients and dL/dV at
sampled points against central differences of the GPU loss (h = 1e-5),
1e-6 relative to the largest component; the GPU forward itself against the
oracle on 4 bands over 5 steps (states 1e-12, loss 1e-11).  The oracle is too
slow for 400 steps x 64 bands at 48^3, so the full-horizon differences use
the GPU loss that the short-horizon comparison pins.

Config 4 (etrs_96): 96^3 grid, 512 bands, 216 atoms.  A captured ETRS step
replayed twice against eager steps, bitwise; the step on 2 bands against the
oracle (1e-12); a 10-step adjoint (checkpoints every 5) against central
differences of the GPU loss on the largest coefficient and one V point.
"""
import sys

import numpy as np
import pytest
import torch

import oracle as O
import inq_amd
from inq_amd import adjoint, real_time, systems
from helpers import relerr

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


OMEGA, NCOEF = 0.1, 32


@pytest.fixture(scope="module")
def config3(device):
    s = systems.diamond_system(1, grid=48)
    psi = systems.random_orbitals(s, 64)
    coef = systems.rng(0, 0, 4).uniform(-0.05, 0.05, size=NCOEF)
    return s, psi, coef


def test_config3_forward_matches_oracle_short_horizon(device, config3):
    s, psi, coef = config3
    say("config 3: 4 bands x 5 steps vs oracle")
    ham = inq_amd.ks_hamiltonian(s, max_bands=64)
    few = np.ascontiguousarray(psi[..., :4])
    prop = adjoint.propagation_adjoint(ham, 4, 5, 0.015, NCOEF, OMEGA, (1.0, 0.0, 0.0),
                                       checkpoint_every=20)
    d = torch.from_numpy(few).to(device)
    out = torch.empty_like(d)
    loss, _ = prop.forward(coef, d, out)
    ref, ref_psi = O.System(s.n, s.lattice, s.vloc, s.tables).td_current(
        few, coef, OMEGA, (1.0, 0.0, 0.0), 0.015, 5)
    err = relerr(out.cpu().numpy(), ref_psi)
    say("psi err %.2e, loss %.15e vs %.15e" % (err, loss, ref))
    assert err < 1e-12
    assert abs(loss - ref) <= 1e-11 * abs(ref)
    prop.close()
    ham.close()


def test_config3_adjoint_matches_finite_differences(device, config3):
    s, psi, coef = config3
    say("config 3: 48^3 x 64, 400 steps, adjoint vs FD on all 32 coefficients")
    ham = inq_amd.ks_hamiltonian(s, max_bands=64)
    prop = adjoint.propagation_adjoint(ham, 64, 400, 0.015, NCOEF, OMEGA, (1.0, 0.0, 0.0),
                                       checkpoint_every=20)
    d = torch.from_numpy(psi).to(device)
    loss, _ = prop.forward(coef, d)
    gc, gV, _ = prop.backward(1.0)
    gV = gV.cpu().numpy().reshape(s.n)
    h = 1e-5
    fd = np.zeros(NCOEF)
    for j in range(NCOEF):
        e = np.zeros(NCOEF)
        e[j] = h
        lp, _ = prop.forward(coef + e, d)
        lm, _ = prop.forward(coef - e, d)
        fd[j] = (lp - lm) / (2 * h)
    scale = np.abs(fd).max()
    err = np.abs(gc - fd).max() / scale
    say("loss %.12e, dL/dc max rel err vs FD %.2e (scale %.3e)" % (loss, err, scale))
    assert err <= 1e-6
    # dL/dV at the largest-gradient point and two seeded points
    rng = np.random.default_rng(48)
    pts = [np.unravel_index(np.argmax(np.abs(gV)), s.n)]
    pts += [tuple(int(rng.integers(0, n)) for n in s.n) for _ in range(2)]
    vscale = np.abs(gV).max()
    v0 = torch.from_numpy(s.vloc.reshape(-1).copy()).to(device)
    for pt in pts:
        flat = int(np.ravel_multi_index(pt, s.n))
        vals = []
        for sgn in (1.0, -1.0):
            v = v0.clone()
            v[flat] += sgn * h
            ham.set_scalar_potential(v)
            vals.append(prop.forward(coef, d)[0])
        fdv = (vals[0] - vals[1]) / (2 * h)
        say("dL/dV", pt, gV[pt], "FD", fdv)
        assert abs(gV[pt] - fdv) <= 1e-6 * vscale
    prop.close()
    ham.close()


@pytest.fixture(scope="module")
def config4(device):
    s = systems.diamond_system(3)
    g = torch.Generator(device=device).manual_seed(4)
    psi = torch.randn(tuple(s.n) + (512,), dtype=torch.complex128, device=device, generator=g)
    psi *= (1.0 / np.sqrt(s.volume))
    coef = systems.rng(0, 0, 4).uniform(-0.05, 0.05, size=NCOEF)
    return s, psi, coef


def test_config4_captured_step_and_oracle(device, config4):
    s, psi, coef = config4
    nst, dt = 512, 0.03
    say("config 4: 96^3 x 512, captured ETRS step vs eager, 2 bands vs oracle")
    torch.cuda.empty_cache()
    ham = inq_amd.ks_hamiltonian(s, max_bands=nst)
    pert = real_time.fourier_series_vector_potential(coef, OMEGA)

    class HostA:
        """The same A(t), evaluated on the host before each replay."""

        def has_uniform_vector_potential(self):
            return True

        def uniform_vector_potential(self, t):
            return pert.uniform_vector_potential(t)

    a = psi.clone()
    b = psi.clone()
    graph = real_time.etrs_graph(ham, a, dt, 4, pert=HostA())
    for step in range(2):
        t = (step + 3) * dt
        graph.launch(t)  # A(t), A(t + dt) written to the device before the replay
        real_time.etrs(ham, b, t, dt, pert=HostA())
        torch.cuda.synchronize()
        assert torch.equal(a, b), step
    graph.close()
    del a
    # one step under A(t) against the oracle on two bands
    c = psi.clone()
    real_time.etrs(ham, c, 0.0, dt, pert)
    lat = s.lattice
    two = np.ascontiguousarray(psi[..., [0, nst - 1]].cpu().numpy())
    ref = O.System(s.n, s.lattice, s.vloc, s.tables).etrs_step(
        two, dt, lat @ pert.uniform_vector_potential(0.0), lat @ pert.uniform_vector_potential(dt))
    err = relerr(c[..., [0, nst - 1]].cpu().numpy(), ref)
    say("config 4 step vs oracle on 2 bands: %.2e" % err)
    assert err < 1e-12
    ham.close()


def test_config4_short_adjoint_matches_finite_differences(device, config4):
    s, psi, coef = config4
    nst, dt = 512, 0.03
    say("config 4: 10-step adjoint (checkpoints every 5) vs FD")
    torch.cuda.empty_cache()
    ham = inq_amd.ks_hamiltonian(s, max_bands=nst)
    prop = adjoint.propagation_adjoint(ham, nst, 10, dt, NCOEF, OMEGA, (1.0, 0.0, 0.0),
                                       checkpoint_every=5, band_slice=32)
    say("propagation memory %.1f GiB" % (prop.memory_bytes / 2**30))
    loss, _ = prop.forward(coef, psi)
    gc, gV, _ = prop.backward(1.0)
    h = 1e-5
    j = int(np.argmax(np.abs(gc)))
    e = np.zeros(NCOEF)
    e[j] = h
    fd = (prop.forward(coef + e, psi)[0] - prop.forward(coef - e, psi)[0]) / (2 * h)
    say("loss %.12e dL/dc_%d %.12e FD %.12e" % (loss, j, gc[j], fd))
    assert abs(gc[j] - fd) <= 1e-6 * np.abs(gc).max()
    gVh = gV.cpu().numpy()
    v0 = torch.from_numpy(s.vloc.reshape(-1).copy()).to(device)

    def fd_along(dv):
        vals = []
        for sgn in (1.0, -1.0):
            ham.set_scalar_potential(v0 + sgn * h * dv)
            vals.append(prop.forward(coef, psi)[0])
        return (vals[0] - vals[1]) / (2 * h), vals

    # dL/dV along a seeded random direction: sum_r dL/dV(r) dV(r) against the
    # central difference of L(V + h dV), 1e-6 relative
    dv = torch.from_numpy(np.random.default_rng(96).standard_normal(s.npoints)).to(device)
    fdd, _ = fd_along(dv)
    dd = float(np.dot(gVh, dv.cpu().numpy()))
    say("dL/dV along a random direction: %.12e FD %.12e" % (dd, fdd))
    assert abs(dd - fdd) <= 1e-6 * abs(fdd)
    # at the largest-gradient point: one of 96^3 points moves a loss of ~1e2
    # by ~1e-9 at h = 1e-5, so the difference carries the rounding of the
    # loss, eps |L| / h, on top of the 1e-6 bound
    flat = int(np.argmax(np.abs(gVh)))
    e = torch.zeros_like(v0)
    e[flat] = 1.0
    fdv, vals = fd_along(e)
    noise = 4 * np.finfo(float).eps * max(abs(x) for x in vals) / h
    say("dL/dV at %d: %.12e FD %.12e (rounding floor %.1e)" % (flat, gVh[flat], fdv, noise))
    assert abs(gVh[flat] - fdv) <= 1e-6 * np.abs(gVh).max() + noise
    ham.set_scalar_potential(v0)
    prop.close()
    ham.close()
