"""real_time.propagate against the oracle: the step loop of
real_time::propagate (propagate.hpp:33-133) with energy::calculate after
every ETRS step and the callback's viewables (step, time, energies, current).

Frozen potential on config 0 (32^3, 8 atoms with projectors) under a
Fourier-series A(t), and the interacting (LDA) propagation on a pencil grid.
Tolerance: 1e-12 relative for states after each step, 1e-11 for the energy
terms and the current (relative to their magnitude).
"""
import sys

import numpy as np
import pytest
import torch

import oracle as O
import inq_amd
from inq_amd import real_time, systems, tddft
from helpers import relerr

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def to_dev(a, device, dtype=np.complex128):
    return torch.from_numpy(np.ascontiguousarray(a, dtype)).to(device)


class Recorder:
    def __init__(self):
        self.rows = []

    def __call__(self, v):
        e = v.energy()
        self.rows.append(dict(iter=v.iter(), time=v.time(), last=v.last_iter(), every2=v.every(2),
                              eig=e.eigenvalues, nl=e.non_local, ext=e.external,
                              hartree=e.hartree, xc=e.xc, total=e.total(),
                              J=v.current(), A=v.uniform_vector_potential(),
                              psi=v.orbitals().cpu().numpy()))


class CountingKick(real_time.fourier_series_vector_potential):
    zero_steps = 0

    def zero_step(self, phi):
        CountingKick.zero_steps += 1


def _oracle_energy(osys, x, kA, occ, V, dV):
    eps, _, nl = osys.energy(x, kA)
    rho = O.density(x, occ)
    return float(occ @ eps), float(occ @ nl), dV * float(np.sum(rho * V))


def test_propagate_frozen_matches_oracle(device):
    s = systems.diamond_system(1)
    nst, nsteps, dt = 8, 3, 0.02
    psi = systems.random_orbitals(s, nst, seed=3)
    occ = np.full(nst, 2.0)
    coef = systems.rng(0, 0, 6).uniform(-0.05, 0.05, size=8)
    pert = CountingKick(coef, 0.1, (0.6, 0.0, 0.8))
    say("propagate frozen 32^3 x", nst, "steps", nsteps)
    ham = inq_amd.ks_hamiltonian(s, max_bands=nst)
    rec = Recorder()
    d = to_dev(psi, device)
    CountingKick.zero_steps = 0
    en = real_time.propagate(ham, d, dt, nsteps, pert=pert, callback=rec, occ=occ)
    assert CountingKick.zero_steps == 1
    assert [r["iter"] for r in rec.rows] == [0, 1, 2, 3]
    assert [r["last"] for r in rec.rows] == [False, False, False, True]
    assert [r["every2"] for r in rec.rows] == [False, False, True, True]
    assert np.allclose([r["time"] for r in rec.rows], [0.0, dt, 2 * dt, 3 * dt], atol=1e-15)
    osys = O.System(s.n, s.lattice, s.vloc, s.tables)
    lat = s.lattice
    dV = s.volume / np.prod(s.n)
    V = s.vloc
    x = psi
    for it, r in enumerate(rec.rows):
        t = it * dt
        kA = lat @ pert.uniform_vector_potential(t)
        if it > 0:
            x = osys.etrs_step(x, dt, lat @ pert.uniform_vector_potential(t - dt), kA)
        eig, nl, ext = _oracle_energy(osys, x, kA, occ, V, dV)
        J = osys.current(x, kA, 2.0, commutator="inq")[1]
        err = relerr(r["psi"], x)
        say("step", it, "psi err %.2e" % err, "eig", r["eig"], eig, "J", r["J"], J)
        assert err < 1e-12
        assert abs(r["eig"] - eig) <= 1e-11 * abs(eig)
        assert abs(r["nl"] - nl) <= 1e-11 * max(abs(nl), 1e-3)
        assert abs(r["ext"] - ext) <= 1e-11 * abs(ext)
        assert np.abs(r["J"] - J).max() <= 1e-11 * np.abs(J).max()
        assert np.allclose(r["A"], pert.uniform_vector_potential(t), rtol=0, atol=1e-15)
    # the energy total combines the terms as energy.hpp's kinetic()/total()
    last = rec.rows[-1]
    assert last["total"] == pytest.approx(en.total(), rel=0, abs=0)
    assert en.kinetic() == pytest.approx(last["eig"] - last["ext"] - last["nl"], rel=1e-14)
    # a restart from step 2 of the same states makes the same last step and no step-0 call
    ham2 = inq_amd.ks_hamiltonian(s, max_bands=nst)
    rec2 = Recorder()
    d2 = to_dev(rec.rows[2]["psi"], device)
    real_time.propagate(ham2, d2, dt, nsteps, pert=pert, start_step=2, callback=rec2, occ=occ)
    assert CountingKick.zero_steps == 1
    assert [r["iter"] for r in rec2.rows] == [3]
    assert np.array_equal(rec2.rows[0]["psi"], rec.rows[3]["psi"])
    ham.close()
    ham2.close()


def test_propagate_interacting_matches_oracle(device):
    lat = np.diag([6.5, 6.5, 6.5])
    s = systems.make_system((12, 14, 12), lat, np.array([[0.3, 0.2, 0.1], [3.4, 3.0, 3.5]]),
                            seed=3, system_id=55)
    nst, nsteps, dt = 4, 2, 0.02
    psi = systems.random_orbitals(s, nst, ecut=6.0)
    occ = np.full(nst, 2.0)
    coef = systems.rng(0, 0, 4).uniform(-0.05, 0.05, size=8)
    pert = real_time.fourier_series_vector_potential(coef, 0.1)
    say("propagate interacting (LDA) 12x14x12 x", nst)
    ham = inq_amd.ks_hamiltonian(s, max_bands=nst)
    rec = Recorder()
    d = to_dev(psi, device)
    vion = to_dev(s.vloc, device, np.float64)
    real_time.propagate(ham, d, dt, nsteps, pert=pert, callback=rec, occ=occ,
                        theory=dict(hartree=True, xc="lda"), vion=vion)
    base = O.System(s.n, s.lattice, s.vloc, s.tables)
    dV = s.volume / np.prod(s.n)
    x = psi
    v, e = base.ks_potential(s.vloc, O.density(x, occ), True, "lda")
    for it, r in enumerate(rec.rows):
        t = it * dt
        kA = lat @ pert.uniform_vector_potential(t)
        if it > 0:
            A0 = lat @ pert.uniform_vector_potential(t - dt)
            x, v, _ = O.System(s.n, s.lattice, v, s.tables).etrs_scf(x, s.vloc, occ, dt, A0, kA,
                                                                      xc="lda")
            v, e = base.ks_potential(s.vloc, O.density(x, occ), True, "lda")
        osys = O.System(s.n, s.lattice, v, s.tables)
        eps, _, nl = osys.energy(x, kA)
        err = relerr(r["psi"], x)
        say("step", it, "psi err %.2e" % err, "eig", r["eig"], float(occ @ eps))
        assert err < 1e-11
        assert abs(r["eig"] - float(occ @ eps)) <= 1e-10 * abs(float(occ @ eps))
        assert abs(r["hartree"] - e["hartree"]) <= 1e-10 * abs(e["hartree"])
        assert abs(r["xc"] - e["exc"]) <= 1e-10 * abs(e["exc"])
        assert abs(r["ext"] - e["external"]) <= 1e-10 * abs(e["external"])
    ham.close()


def test_taylor_chain_equals_one_hpsi_per_order(device, tmp_path):
    """The chunk-by-chunk Taylor chain (each order's last sweep transforms its
    output forward for the next order) against one Hpsi per order
    (INQAMD_TAYLOR_CHAIN=0), bitwise: exponential_in_place, 2-for-1, ETRS
    steps at 48^3 and 64^3 (band groups), a captured step and the interacting
    ETRS.  The library reads the switch once per process, so each form runs
    in a child process (tests/taylor_chain_dump.py)."""
    import os
    import subprocess
    script = os.path.join(os.path.dirname(os.path.abspath(__file__)), "taylor_chain_dump.py")
    res = {}
    for chain in ("1", "0"):
        say("taylor chain", chain)
        e = dict(os.environ)
        e["INQAMD_TAYLOR_CHAIN"] = chain
        path = str(tmp_path / f"chain{chain}.npz")
        r = subprocess.run([sys.executable, "-u", script, path], env=e, capture_output=True,
                           text=True, timeout=240)
        say(r.stdout[-1500:], r.stderr[-1500:])
        assert r.returncode == 0 and "CHAIN-DUMP-OK" in r.stdout
        res[chain] = np.load(path)
    for k in res["0"].files:
        same = np.array_equal(res["1"][k], res["0"][k])
        say(k, "bitwise equal" if same else "DIFFERS")
        assert same, k
