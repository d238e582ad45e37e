"""Hpsi of many independent systems in one call (uneven batches).

    batch = hamiltonian_batch([ks_hamiltonian(s, nst) for s, nst in ...])
    outs = batch([phi_0, phi_1, ...])            # one Hpsi per system

Each phase of Hpsi is one persistent work-queue launch over the whole batch
(include/inq_amd.h, inqamd_batch_*); results match each system's own
ks_hamiltonian(phi) to rounding.  `autotune` picks the tiling knobs by
timing the batch on the device.
"""
import ctypes
import itertools
import json
import os

import torch

from . import _lib
from .hamiltonian import _orbital_args, _stream_ptr


class hamiltonian_batch:
    def __init__(self, hams):
        self.hams = list(hams)
        L = _lib.lib()
        self._lib = L
        arr = (ctypes.c_void_p * len(self.hams))(*[h.handle.h.value for h in self.hams])
        b = ctypes.c_void_p()
        rc = L.inqamd_batch_create(arr, len(self.hams), _stream_ptr(), ctypes.byref(b))
        if rc != _lib.OK:
            raise _lib.InqAmdError(rc, (L.inqamd_batch_last_error(None) or b"").decode())
        self._b = b
        self.tuning = dict(chunk_bytes=128 << 20, ty_main=8, narrow_remainders=True,
                           group_by_grid=False, nstreams=2)

    def _check(self, rc):
        if rc != _lib.OK:
            raise _lib.InqAmdError(rc, (self._lib.inqamd_batch_last_error(self._b) or b"").decode())

    def set_tuning(self, chunk_bytes=128 << 20, ty_main=8, narrow_remainders=True,
                   group_by_grid=False, nstreams=2):
        self._check(self._lib.inqamd_batch_set_tuning(self._b, int(chunk_bytes), int(ty_main),
                                                      int(bool(narrow_remainders)),
                                                      int(bool(group_by_grid)), int(nstreams)))
        self.tuning = dict(chunk_bytes=int(chunk_bytes), ty_main=int(ty_main),
                           narrow_remainders=bool(narrow_remainders),
                           group_by_grid=bool(group_by_grid), nstreams=int(nstreams))

    def __call__(self, phis, outs=None):
        if len(phis) != len(self.hams):
            raise ValueError(f"{len(phis)} orbital sets for {len(self.hams)} systems")
        if outs is None:
            outs = [torch.empty_like(p) for p in phis]
        n = len(phis)
        ps, qs = (ctypes.c_void_p * n)(), (ctypes.c_void_p * n)()
        nst, ld = (ctypes.c_int * n)(), (ctypes.c_int * n)()
        for i, (p, q) in enumerate(zip(phis, outs)):
            pp, ns, l = _orbital_args(p)
            qq, ns2, _ = _orbital_args(q)
            if ns2 != ns:
                raise ValueError("output band count differs from input")
            ps[i], qs[i], nst[i], ld[i] = pp, qq, ns, l
        self._check(self._lib.inqamd_batch_set_stream(self._b, _stream_ptr()))
        self._check(self._lib.inqamd_batch_apply(self._b, ps, qs, nst, ld))
        return outs

    def info(self):
        v = [ctypes.c_int(0) for _ in range(4)]
        self._check(self._lib.inqamd_batch_info(self._b, *[ctypes.byref(x) for x in v]))
        return dict(chunks=v[0].value, launches=v[1].value, pieces=v[2].value,
                    fallback=v[3].value)

    def close(self):
        if getattr(self, "_b", None):
            self._lib.inqamd_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_TUNING_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def committed_tuning(name):
    """The committed set_tuning arguments for a named batch (tuning.json)."""
    with open(_TUNING_PATH) as f:
        t = json.load(f)["batch"][name]
    return {k: t[k] for k in ("chunk_bytes", "ty_main", "narrow_remainders", "group_by_grid",
                              "nstreams")}


def _time(fn, reps):
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fn()
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def autotune(batch, phis, outs, reps=5, chunk_mib=(32, 64, 128, 256), ty_main=(8, 16),
             narrow=(True, False), group=(False, True), nstreams=(1, 2, 4)):
    """Time every combination of the tiling knobs on the device and keep the
    fastest; returns (best settings, {settings tuple: ms})."""
    results = {}
    for c, t, nr, g, ns in itertools.product(chunk_mib, ty_main, narrow, group, nstreams):
        batch.set_tuning(c << 20, t, nr, g, ns)
        results[(c, t, nr, g, ns)] = _time(lambda: batch(phis, outs), reps)
    best = min(results, key=results.get)
    batch.set_tuning(best[0] << 20, best[1], best[2], best[3], best[4])
    return dict(batch.tuning), results
