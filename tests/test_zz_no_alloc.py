"""inqamd_band_energies_noalloc (the allocation-free band energies, partial
sums in the library's fixed device scratch) against inqamd_band_energies
(partial sums in the handle's workspace), both through the C ABI: equal bit
for bit, on two calls in a row whose second has a smaller input, on two
handles; sizes beyond the scratch are refused with the limit in the message.
"""
import ctypes
import sys

import numpy as np
import pytest
import torch

import inq_amd
from inq_amd import _lib, systems

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _energies(L, h, name, hpsi, psi, nst, ld, stream):
    eps = torch.full((nst,), float("nan"), dtype=torch.float64, device=psi.device)
    _lib.check(L.inqamd_set_stream(h, stream), h)
    rc = getattr(L, name)(h, ctypes.c_void_p(hpsi.data_ptr()), ctypes.c_void_p(psi.data_ptr()),
                          ctypes.c_void_p(eps.data_ptr()), nst, ld)
    _lib.check(rc, h)
    torch.cuda.synchronize()
    return eps.cpu().numpy()


def test_noalloc_band_energies_equal_the_workspace_path(device):
    L = _lib.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    cases = [(2, 96, "64^3 x 96 bands"), (2, 40, "64^3 x 40 bands (smaller, same handle)"),
             (1, 16, "32^3 x 16 bands (smaller grid, another handle)")]
    hams = {}
    for sc, nst, what in cases:
        say("band energies, no-alloc vs workspace:", what)
        if sc not in hams:
            hams[sc] = inq_amd.ks_hamiltonian(systems.diamond_system(sc), max_bands=96)
        ham = hams[sc]
        s = ham.system
        psi = torch.from_numpy(systems.random_orbitals(s, nst, seed=7)).to(device)
        hpsi = ham(psi)
        torch.cuda.synchronize()
        ref = _energies(L, ham._h.h, "inqamd_band_energies", hpsi, psi, nst, nst, stream)
        got = _energies(L, ham._h.h, "inqamd_band_energies_noalloc", hpsi, psi, nst, nst, stream)
        assert np.all(np.isfinite(ref))
        say("max |eps| %.6e, max |diff| %.1e" % (np.abs(ref).max(), np.abs(got - ref).max()))
        assert np.array_equal(got, ref)
    for ham in hams.values():
        ham.close()


def test_noalloc_band_energies_refuse_sizes_beyond_the_scratch(device):
    L = _lib.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    s = systems.diamond_system(1)  # 32^3: 128 point chunks, at most 4096 bands
    ham = inq_amd.ks_hamiltonian(s, max_bands=8)
    h = ham._h.h
    small = torch.zeros(8, dtype=torch.complex128, device=device)
    eps = torch.zeros(8, dtype=torch.float64, device=device)
    _lib.check(L.inqamd_set_stream(h, stream), h)
    # refused before any launch: the buffers are never read
    rc = L.inqamd_band_energies_noalloc(h, ctypes.c_void_p(small.data_ptr()),
                                        ctypes.c_void_p(small.data_ptr()),
                                        ctypes.c_void_p(eps.data_ptr()), 4097, 4097)
    msg = L.inqamd_last_error(h)
    say("4097 bands at 32^3:", rc, msg)
    assert rc == _lib.ERR_ARGUMENT
    assert b"at most 4096 bands" in msg
    ham.close()
