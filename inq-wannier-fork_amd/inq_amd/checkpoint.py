"""Checkpoints in INQ's on-disk format.

Orbitals: operations::io::save/load of an orbital_set (src/operations/io.hpp:
163-197) through the library (one raw complex128 file per band,
"%010d_%010d").  Around them the layout of systems::electrons::save/try_load
(src/systems/electrons.hpp:323-370): kpin%010d/states, kpin%010d/occupations/
array (raw doubles) and spin_density/%010d (raw doubles per component); and
the real-time checkpoint of real_time::results (src/real_time/results.hpp:
53-112): .inq/default_checkpoint/real-time/{observables,orbitals} and a
`type` file written last.  Text values follow utils::save_value/
save_container (src/utils/load_save.hpp: 25 significant digits, one value
per line, vector3 as tab-separated components).
"""
import os

import numpy as np
import torch

from .hamiltonian import _orbital_args


def _num(v):
    return "%010d" % v  # utils::num_to_str (src/utils/num_str.hpp:20-24)


def save_states(ham, psi, dirname, first_band=0, spinor=0):
    """operations::io::save(dirname, orbital_set) for one spinor component."""
    p, nst, ld = _orbital_args(psi)
    ham.handle.call("inqamd_save_states", p, nst, ld, int(first_band), int(spinor),
                    os.fsencode(dirname))


def load_states(ham, psi, dirname, first_band=0, spinor=0):
    """operations::io::load into psi (in place); raises if a file is missing."""
    p, nst, ld = _orbital_args(psi)
    ham.handle.call("inqamd_load_states", p, nst, ld, int(first_band), int(spinor),
                    os.fsencode(dirname))
    return psi


def _save_raw(dirname, arr):
    os.makedirs(dirname, exist_ok=True)
    np.ascontiguousarray(arr, np.float64).tofile(os.path.join(dirname, "array"))


def save_electrons(dirname, ham, psi, occupations, spin_density=None, kpin=0):
    """systems::electrons::save for one k-point/spin set (kpin index `kpin`)."""
    base = os.path.join(dirname, "kpin" + _num(kpin))
    save_states(ham, psi, os.path.join(base, "states"))
    _save_raw(os.path.join(base, "occupations"), np.broadcast_to(
        np.asarray(occupations, np.float64), (psi.shape[-1],)))
    if spin_density is not None:
        rho = spin_density.detach().cpu().numpy() if torch.is_tensor(spin_density) \
            else np.asarray(spin_density, np.float64)
        rho = rho.reshape((-1,) + tuple(ham.n)) if rho.ndim == 4 else rho.reshape((1,) + tuple(ham.n))
        os.makedirs(os.path.join(dirname, "spin_density"), exist_ok=True)
        # field_set save: one file per component (io.hpp:131-141)
        for c in range(rho.shape[0]):
            np.ascontiguousarray(rho[c], np.float64).tofile(
                os.path.join(dirname, "spin_density", _num(c)))


def load_electrons(dirname, ham, psi, kpin=0, ncomponents=1):
    """systems::electrons::try_load: fills psi; returns (occupations, spin
    density [ncomponents, nx, ny, nz] or None if absent)."""
    base = os.path.join(dirname, "kpin" + _num(kpin))
    load_states(ham, psi, os.path.join(base, "states"))
    occ = np.fromfile(os.path.join(base, "occupations", "array"), np.float64)
    if occ.size != psi.shape[-1]:
        raise ValueError("occupations file holds %d values for %d bands" % (occ.size, psi.shape[-1]))
    rho = None
    sd = os.path.join(dirname, "spin_density")
    if os.path.isdir(sd):
        N = int(np.prod(ham.n))
        rho = np.stack([np.fromfile(os.path.join(sd, _num(c)), np.float64, count=N).reshape(ham.n)
                        for c in range(ncomponents)])
    return occ, rho


def _fmt(v):
    # std::ostream with precision(25): the %g form with 25 significant digits
    return "%.25g" % v


def _write_value(path, v):
    with open(path, "w") as f:
        f.write(_fmt(v) + "\n" if isinstance(v, float) else "%s\n" % v)


def _write_container(path, values):
    if len(values) == 0:
        if os.path.exists(path):
            os.remove(path)
        return
    with open(path, "w") as f:
        for v in values:
            if np.ndim(v) == 1:
                f.write("\t".join(_fmt(float(x)) for x in v) + "\n")
            else:
                f.write(_fmt(float(v)) + "\n")


def save_real_time_results(dirname, total_steps, total_time, time, total_energy, dipole=(),
                           current=()):
    """real_time::results::save (results.hpp:70-80)."""
    os.makedirs(dirname, exist_ok=True)
    _write_value(os.path.join(dirname, "total_steps"), int(total_steps))
    _write_value(os.path.join(dirname, "total_time"), float(total_time))
    _write_container(os.path.join(dirname, "time"), list(time))
    _write_container(os.path.join(dirname, "total_energy"), list(total_energy))
    _write_container(os.path.join(dirname, "dipole"), list(dipole))
    _write_container(os.path.join(dirname, "current"), list(current))


def load_real_time_results(dirname):
    """real_time::results::load (results.hpp:82-107) as a dict."""
    def rd(name):
        p = os.path.join(dirname, name)
        if not os.path.exists(p):
            return []
        return [list(map(float, line.split())) for line in open(p) if line.strip()]
    res = {"total_steps": int(open(os.path.join(dirname, "total_steps")).read().split()[0]),
           "total_time": float(open(os.path.join(dirname, "total_time")).read().split()[0])}
    res["time"] = [v[0] for v in rd("time")]
    res["total_energy"] = [v[0] for v in rd("total_energy")]
    res["dipole"] = rd("dipole")
    res["current"] = rd("current")
    return res


def _write_checkpoint(root, results, write_orbitals):
    # the type file is removed first and written last (results.hpp:60-65)
    os.makedirs(root, exist_ok=True)
    tfile = os.path.join(root, "type")
    if os.path.exists(tfile):
        os.remove(tfile)
    save_real_time_results(os.path.join(root, "real-time", "observables"), **results)
    write_orbitals(os.path.join(root, "real-time", "orbitals"))
    _write_value(tfile, "real-time")


def save_real_time_checkpoint(ham, psi, occupations, spin_density, results,
                              root=".inq/default_checkpoint"):
    """The checkpoint real_time::results writes every 500 steps
    (results.hpp:53-66): the type file is removed first and written last, so
    a partial checkpoint has none."""
    _write_checkpoint(root, results, lambda d: save_electrons(d, ham, psi, occupations,
                                                              spin_density))


def save_real_time_checkpoint_host(psi, occupations, results, root=".inq/default_checkpoint",
                                   kpin=0):
    """The same checkpoint from host (numpy) orbitals, without the density."""
    def orbitals(d):
        base = os.path.join(d, "kpin" + _num(kpin))
        states_from_host(os.path.join(base, "states"), psi)
        _save_raw(os.path.join(base, "occupations"), np.broadcast_to(
            np.asarray(occupations, np.float64), (np.shape(psi)[-1],)))
    _write_checkpoint(root, results, orbitals)


class CheckpointError(RuntimeError):
    """A checkpoint that cannot be resumed (the messages of run.hpp:121-151)."""


def checkpoint_type(root=".inq/default_checkpoint"):
    """utils::load_optional(root/type): the checkpoint type or None."""
    p = os.path.join(root, "type")
    if not os.path.exists(p):
        return None
    return open(p).read().split("\n")[0]


def resume_real_time(root=".inq/default_checkpoint"):
    """The checks of run::resume (interface/run.hpp:121-151) before a
    real-time run continues from a checkpoint: the `type` file (written last,
    so a partially written checkpoint has none) must exist and read
    "real-time".  Returns (observables dict, orbitals directory); the
    propagation then continues with start_step = total_steps."""
    t = checkpoint_type(root)
    if t is None:
        raise CheckpointError("Cannot resume run, a checkpoint was not found.")
    if t != "real-time":
        raise CheckpointError("Unknown checkpoint type '%s'." % t)
    orb = os.path.join(root, "real-time", "orbitals")
    if not os.path.isdir(os.path.join(orb, "kpin" + _num(0), "states")):
        raise CheckpointError("Cannot load the restart electron orbitals.\n"
                              " The checkpoint must be corrupt")
    return load_real_time_results(os.path.join(root, "real-time", "observables")), orb


def states_to_host(dirname, n, nst, first_band=0, spinor=0):
    """Read the per-band orbital files of operations::io::save into a numpy
    array [nx, ny, nz, nst] (host side; the library's load_states does the
    same on the device)."""
    N = int(np.prod(n))
    out = np.empty((N, nst), np.complex128)
    for b in range(nst):
        p = os.path.join(dirname, "%s_%s" % (_num(first_band + b), _num(spinor)))
        if not os.path.exists(p):
            raise CheckpointError("cannot open restart file " + p)
        raw = np.fromfile(p, np.complex128)
        if raw.size != N:
            raise CheckpointError("restart file %s holds %d points, expected %d" % (p, raw.size, N))
        out[:, b] = raw
    return out.reshape(tuple(n) + (nst,))


def states_from_host(dirname, psi, first_band=0, spinor=0):
    """Write psi [.., nst] (numpy) as operations::io::save does: one raw
    complex128 file per band in grid order."""
    os.makedirs(dirname, exist_ok=True)
    a = np.ascontiguousarray(psi, np.complex128)
    a = a.reshape(-1, a.shape[-1])
    for b in range(a.shape[-1]):
        np.ascontiguousarray(a[:, b]).tofile(
            os.path.join(dirname, "%s_%s" % (_num(first_band + b), _num(spinor))))


class real_time_results:
    """real_time::results (src/real_time/results.hpp:24-112): the propagation
    callback that records time, total energy and, if asked, the dipole and
    the current of every step, and writes the real-time checkpoint every
    `every` steps (500 in INQ) through `save_checkpoint(viewables, results)`.
    """

    def __init__(self, dipole=False, current=False, every=500, save_checkpoint=None):
        self.total_steps = 0
        self.total_time = 0.0
        self.time, self.total_energy, self.dipole, self.current = [], [], [], []
        self.want_dipole, self.want_current = bool(dipole), bool(current)
        self.every_steps = int(every)
        self.save_checkpoint = save_checkpoint

    @classmethod
    def load(cls, dirname, **kw):
        r = cls(**kw)
        d = load_real_time_results(dirname)
        r.total_steps, r.total_time = d["total_steps"], d["total_time"]
        r.time, r.total_energy = list(d["time"]), list(d["total_energy"])
        r.dipole, r.current = [tuple(v) for v in d["dipole"]], [tuple(v) for v in d["current"]]
        return r

    def as_dict(self):
        return dict(total_steps=self.total_steps, total_time=self.total_time, time=self.time,
                    total_energy=self.total_energy, dipole=self.dipole, current=self.current)

    def __call__(self, v):
        self.total_steps = v.iter()
        self.total_time = v.time()
        self.time.append(v.time())
        self.total_energy.append(v.energy().total())
        if self.want_dipole:
            self.dipole.append(tuple(float(x) for x in v.dipole()))
        if self.want_current:
            self.current.append(tuple(float(x) for x in v.current()))
        if self.total_steps + 1 != len(self.time):
            raise RuntimeError("real_time_results: %d steps but %d recorded times"
                               % (self.total_steps, len(self.time)))
        if self.save_checkpoint is not None and v.every(self.every_steps):
            self.save_checkpoint(v, self)
