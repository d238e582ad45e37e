"""GPU tests of the checkpoint files: orbital files in INQ's format (one raw
complex128 file per band, "%010d_%010d", grid order) read back with numpy,
bitwise round trips through the library, band offsets and spinor suffixes,
a missing file, and the electrons / real-time checkpoint layout."""
import os
import sys

import numpy as np
import pytest
import torch

import inq_amd
from inq_amd import checkpoint, systems

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def cfg(device):
    s = systems.make_system((12, 10, 14), np.diag([5.0, 4.5, 6.0]), np.array([[1.0, 1.0, 1.0]]),
                            seed=2, system_id=61)
    psi = systems.random_orbitals(s, 6, ecut=5.0)
    return s, psi


def test_states_files_and_round_trip(device, cfg, tmp_path):
    s, psi = cfg
    say("states files")
    ham = inq_amd.ks_hamiltonian(s, max_bands=6)
    d = torch.from_numpy(psi).to(device)
    out = str(tmp_path / "states")
    checkpoint.save_states(ham, d, out, first_band=10, spinor=1)
    names = sorted(os.listdir(out))
    assert names == ["%010d_%010d" % (10 + b, 1) for b in range(6)]
    N = int(np.prod(s.n))
    for b in range(6):
        raw = np.fromfile(os.path.join(out, names[b]), np.complex128)
        assert raw.size == N
        assert np.array_equal(raw, psi[..., b].ravel())
    back = torch.zeros_like(d)
    checkpoint.load_states(ham, back, out, first_band=10, spinor=1)
    assert np.array_equal(back.cpu().numpy().view(np.uint64), psi.view(np.uint64))
    # a band range past the saved files fails like INQ's load (returns false)
    with pytest.raises(Exception, match="cannot open restart file"):
        checkpoint.load_states(ham, back, out, first_band=11, spinor=1)
    ham.close()


def test_electrons_and_real_time_checkpoint(device, cfg, tmp_path):
    s, psi = cfg
    say("electrons checkpoint")
    ham = inq_amd.ks_hamiltonian(s, max_bands=6)
    d = torch.from_numpy(psi).to(device)
    occ = np.array([2.0, 2.0, 1.5, 1.0, 0.5, 0.0])
    rho = inq_amd.tddft.density(ham, d, occ)
    root = str(tmp_path / "ck")
    res = dict(total_steps=1, total_time=0.05, time=[0.0, 0.05], total_energy=[-3.0, -3.0],
               current=[(0.0, 0.0, 0.0), (0.1, 0.2, 0.3)])
    checkpoint.save_real_time_checkpoint(ham, d, occ, rho, res, root=root)
    assert open(os.path.join(root, "type")).read() == "real-time\n"
    orb = os.path.join(root, "real-time", "orbitals")
    assert os.path.isfile(os.path.join(orb, "kpin0000000000", "states", "0000000005_0000000000"))
    assert np.array_equal(np.fromfile(os.path.join(orb, "kpin0000000000", "occupations", "array")), occ)
    raw_rho = np.fromfile(os.path.join(orb, "spin_density", "0000000000"), np.float64)
    assert np.array_equal(raw_rho, rho.cpu().numpy().ravel())
    back = torch.zeros_like(d)
    occ2, rho2 = checkpoint.load_electrons(orb, ham, back)
    assert np.array_equal(occ2, occ) and np.array_equal(rho2[0], rho.cpu().numpy())
    assert np.array_equal(back.cpu().numpy(), psi)
    r = checkpoint.load_real_time_results(os.path.join(root, "real-time", "observables"))
    assert r["total_steps"] == 1 and r["current"][1] == [0.1, 0.2, 0.3]
    ham.close()
