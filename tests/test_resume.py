"""Resume of a real-time run from INQ's checkpoint (CPU only, with the
oracle as the propagator): the `type` marker is removed first and written
last (real_time/results.hpp:53-65), so a checkpoint interrupted while being
written -- or one whose marker is gone -- is refused with run::resume's
messages (interface/run.hpp:121-151); a run resumed from a checkpoint
reproduces the uninterrupted run bitwise (states and recorded observables)."""
import os

import numpy as np
import pytest

import oracle as O
from inq_amd import checkpoint, real_time, systems

DT, STEPS, SPLIT = 0.02, 6, 3


@pytest.fixture(scope="module")
def case():
    s = systems.make_system((10, 12, 10), np.diag([5.5, 6.0, 5.5]),
                            np.array([[0.2, 0.1, -0.3], [2.6, 2.9, 2.4]]), seed=6, system_id=71)
    psi = systems.random_orbitals(s, 2, ecut=5.0)
    pert = real_time.fourier_series_vector_potential([0.03, -0.02], 0.4, (1.0, 0.0, 0.0))
    return s, psi, pert, O.System(s.n, s.lattice, s.vloc, s.tables)


class _View:
    """The parts of real_time.viewables a results recorder reads."""

    def __init__(self, it, last, t, total):
        self._it, self._last, self._t, self._total = it, last, t, total

    def iter(self):
        return self._it

    def last_iter(self):
        return self._last

    def every(self, n):
        return self._it != 0 and (self._it % n == 0 or self._last)

    def time(self):
        return self._t

    def energy(self):
        total = self._total

        class E:
            def total(self):
                return total
        return E()


def _propagate(case, psi, start, res):
    """propagate's loop on the oracle: ETRS, then the energy, then the callback."""
    s, _, pert, osys = case
    lat = s.lattice
    occ = np.full(psi.shape[-1], 2.0)

    def total(x, t):
        eps, _, _ = osys.energy(x, lat @ pert.uniform_vector_potential(t))
        return float(occ @ eps)

    if start == 0:
        res(_View(0, False, 0.0, total(psi, 0.0)))
    for n in range(start, STEPS):
        t = n * DT
        psi = osys.etrs_step(psi, DT, lat @ pert.uniform_vector_potential(t),
                             lat @ pert.uniform_vector_potential(t + DT))
        res(_View(n + 1, n == STEPS - 1, (n + 1) * DT, total(psi, t + DT)))
    return psi


def test_resumed_run_is_bitwise_the_uninterrupted_one(case, tmp_path):
    s, psi0, _, _ = case
    straight = checkpoint.real_time_results()
    psi_a = _propagate(case, psi0, 0, straight)
    # the same run, checkpointed every SPLIT steps and stopped after the first
    root = str(tmp_path / "ck")
    saved = {}

    def save(v, res):
        checkpoint.save_real_time_checkpoint_host(saved["psi"], 2.0, res.as_dict(), root=root)
        raise KeyboardInterrupt  # the run ends right after its first checkpoint

    first = checkpoint.real_time_results(every=SPLIT, save_checkpoint=save)

    def run_until_checkpoint():
        lat = s.lattice
        osys, pert = case[3], case[2]
        occ = np.full(psi0.shape[-1], 2.0)
        x = psi0
        e0 = osys.energy(x, lat @ pert.uniform_vector_potential(0.0))[0]
        first(_View(0, False, 0.0, float(occ @ e0)))
        for n in range(STEPS):
            t = n * DT
            x = osys.etrs_step(x, DT, lat @ pert.uniform_vector_potential(t),
                               lat @ pert.uniform_vector_potential(t + DT))
            saved["psi"] = x
            eps = osys.energy(x, lat @ pert.uniform_vector_potential(t + DT))[0]
            first(_View(n + 1, n == STEPS - 1, (n + 1) * DT, float(occ @ eps)))

    with pytest.raises(KeyboardInterrupt):
        run_until_checkpoint()
    assert checkpoint.checkpoint_type(root) == "real-time"
    obs, orb = checkpoint.resume_real_time(root)
    assert obs["total_steps"] == SPLIT
    resumed = checkpoint.real_time_results.load(os.path.join(root, "real-time", "observables"))
    psi_r = checkpoint.states_to_host(os.path.join(orb, "kpin0000000000", "states"), s.n,
                                      psi0.shape[-1])
    psi_b = _propagate(case, psi_r, resumed.total_steps, resumed)
    assert np.array_equal(psi_a.view(np.uint64), psi_b.view(np.uint64))
    assert resumed.total_steps == straight.total_steps == STEPS
    assert resumed.time == straight.time
    assert resumed.total_energy == straight.total_energy


def test_checkpoint_without_marker_is_refused(case, tmp_path):
    s, psi0, _, _ = case
    root = str(tmp_path / "ck")
    res = dict(total_steps=2, total_time=2 * DT, time=[0.0, DT, 2 * DT],
               total_energy=[-1.0, -1.0, -1.0])
    checkpoint.save_real_time_checkpoint_host(psi0, 2.0, res, root=root)
    assert checkpoint.resume_real_time(root)[0]["total_steps"] == 2
    os.remove(os.path.join(root, "type"))
    with pytest.raises(checkpoint.CheckpointError, match="a checkpoint was not found"):
        checkpoint.resume_real_time(root)
    with open(os.path.join(root, "type"), "w") as f:
        f.write("ground-state\n")
    with pytest.raises(checkpoint.CheckpointError, match="Unknown checkpoint type"):
        checkpoint.resume_real_time(root)


def test_interrupted_checkpoint_write_leaves_no_marker(case, tmp_path, monkeypatch):
    """A failure while the orbitals are written (after the old marker was
    removed) leaves a checkpoint that resume refuses."""
    s, psi0, _, _ = case
    root = str(tmp_path / "ck")
    res = dict(total_steps=1, total_time=DT, time=[0.0, DT], total_energy=[-1.0, -1.0])
    checkpoint.save_real_time_checkpoint_host(psi0, 2.0, res, root=root)
    assert checkpoint.checkpoint_type(root) == "real-time"

    def fail(*a, **k):
        raise OSError("disk full")

    monkeypatch.setattr(checkpoint, "states_from_host", fail)
    res2 = dict(res, total_steps=2, total_time=2 * DT)
    with pytest.raises(OSError, match="disk full"):
        checkpoint.save_real_time_checkpoint_host(psi0, 2.0, res2, root=root)
    assert checkpoint.checkpoint_type(root) is None
    with pytest.raises(checkpoint.CheckpointError, match="not found"):
        checkpoint.resume_real_time(root)
