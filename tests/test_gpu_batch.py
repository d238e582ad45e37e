"""GPU tests of the uneven-batch Hpsi.

The batch runs the same tile bodies and projector blocks as the one-system
Hpsi; its output must match each system's own ks_hamiltonian(phi) to 1e-14
relative (different kernel instantiations may contract multiply-adds
differently, so agreement is not bitwise) under every tiling choice, and
every member must match the oracle to 1e-12 relative.
"""
import sys

import numpy as np
import pytest
import torch

import oracle as O
from inq_amd import batch, hamiltonian, systems
from helpers import relerr

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def to_dev(a, device):
    return torch.from_numpy(np.ascontiguousarray(a, np.complex128)).to(device)


# (grid, bands, atoms): band remainders of every size class, many overlapping
# spheres in a small cell, a non-cubic grid and a grid outside the batch
# kernels' table (both run the one-system Hpsi inside the call)
SMALL = [((24, 24, 24), 3, 2), ((30, 30, 30), 13, 5), ((36, 36, 36), 8, 3),
         ((24, 24, 24), 21, 40), ((40, 40, 40), 5, 1), ((16, 18, 20), 3, 2), ((80, 80, 80), 2, 1),
         ((32, 32, 32), 17, 0)]


@pytest.fixture(scope="module")
def small_batch(device):
    sysl, hams, phis = [], [], []
    for i, (n, nst, na) in enumerate(SMALL):
        h = 10.26 / 32
        lat = np.diag([n[0] * h, n[1] * h, n[2] * h])
        frac = systems.rng(5, 300 + i, 2).uniform(0, 1, size=(na, 3))
        s = systems.make_system(n, lat, frac @ lat, seed=5, system_id=300 + i,
                                with_projectors=na > 0)
        psi = systems.random_orbitals(s, nst, seed=5, system_id=300 + i, ecut=8.0)
        sysl.append((s, psi))
        hams.append(hamiltonian.ks_hamiltonian(s, nst))
        phis.append(to_dev(psi, device))
    refs = [h(p) for h, p in zip(hams, phis)]
    torch.cuda.synchronize()
    refs = [r.cpu().numpy() for r in refs]
    yield sysl, hams, phis, refs
    for h in hams:
        h.close()


@pytest.mark.parametrize("tuning", [
    dict(),
    dict(group_by_grid=True),
    dict(ty_main=16, narrow_remainders=False),
    dict(ty_main=32, group_by_grid=True),
    dict(chunk_bytes=1 << 20),
    dict(chunk_bytes=1 << 20, group_by_grid=True, narrow_remainders=False),
    dict(chunk_bytes=1 << 20, group_by_grid=True, nstreams=4),
    dict(nstreams=1),
])
def test_batch_matches_per_system(device, small_batch, tuning):
    say("batch small", tuning)
    sysl, hams, phis, refs = small_batch
    b = batch.hamiltonian_batch(hams)
    b.set_tuning(**tuning)
    outs = b(phis)
    torch.cuda.synchronize()
    info = b.info()
    assert info["fallback"] == 2
    errs = [relerr(o.cpu().numpy(), r) for o, r in zip(outs, refs)]
    say("batch vs per-system rel err", ["%.1e" % e for e in errs], info)
    assert max(errs) < 1e-14, (errs, tuning)
    # a second call reuses the work tables
    outs2 = b(phis, outs)
    torch.cuda.synchronize()
    for o, r in zip(outs2, refs):
        assert relerr(o.cpu().numpy(), r) < 1e-14
    b.close()


def test_batch_member_matches_oracle(device, small_batch):
    say("batch member vs oracle")
    sysl, hams, phis, _ = small_batch
    b = batch.hamiltonian_batch(hams)
    outs = b(phis)
    torch.cuda.synchronize()
    for i, (s, psi) in enumerate(sysl):
        ref = O.System(s.n, s.lattice, s.vloc, s.tables).hamiltonian(psi)
        err = relerr(outs[i].cpu().numpy(), ref)
        say("member", i, s.n, psi.shape[-1], "rel err vs oracle %.2e" % err)
        assert err < 1e-12, i
    b.close()


def test_uneven_batch_config_matches_per_system(device):
    """The full 24-system uneven batch: batched output against per-system
    Hpsi, and the smallest system against the oracle."""
    lay = systems.uneven_batch_layout()
    say("uneven batch config", len(lay), "systems")
    hams, phis, sysl = [], [], []
    for i, (L, nst, na) in enumerate(lay):
        s = systems.uneven_batch_system(i, L, na)
        psi = systems.random_orbitals(s, nst, seed=0, system_id=100 + i)
        sysl.append((s, psi))
        hams.append(hamiltonian.ks_hamiltonian(s, nst))
        phis.append(to_dev(psi, device))
    refs = [h(p) for h, p in zip(hams, phis)]
    b = batch.hamiltonian_batch(hams)
    outs = b(phis)
    torch.cuda.synchronize()
    errs = [float((torch.linalg.vector_norm(o - r) / torch.linalg.vector_norm(r)).item())
            for o, r in zip(outs, refs)]
    say("uneven batch vs per-system max rel err %.2e" % max(errs))
    assert max(errs) < 1e-14
    k = int(np.argmin([L ** 3 * nst for L, nst, _ in lay]))
    s, psi = sysl[k]
    ref = O.System(s.n, s.lattice, s.vloc, s.tables).hamiltonian(psi)
    assert relerr(outs[k].cpu().numpy(), ref) < 1e-12
    b.close()
    for h in hams:
        h.close()
