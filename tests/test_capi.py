"""The C ABI library loads and exports every symbol include/inq_amd.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re

import pytest

from inq_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "inq_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(inqamd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_binding_set():
    names = {s[0] for s in _lib.SIGNATURES}
    assert set(declared_symbols()) == names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_version_string():
    assert b"gfx950" in _lib.lib().inqamd_version()


def test_create_rejects_bad_arguments_without_touching_the_gpu():
    L = _lib.lib()
    h = ctypes.c_void_p()
    rc = L.inqamd_hamiltonian_create(None, None, ctypes.byref(h))
    assert rc == _lib.ERR_ARGUMENT and not h.value
    d = _lib.SystemDesc()
    d.n = (ctypes.c_int * 3)(8, 8, 8)
    d.lattice = (ctypes.c_double * 9)(5, 0, 0, 0, 5, 0, 0, 0, 5)
    d.max_bands = 0
    rc = L.inqamd_hamiltonian_create(ctypes.byref(d), None, ctypes.byref(h))
    assert rc == _lib.ERR_ARGUMENT
    assert b"max_bands" in L.inqamd_last_error(None)
    import numpy as np
    v = np.zeros(37 * 64)
    d.vloc = v.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    d.max_bands = 4
    d.n = (ctypes.c_int * 3)(37, 8, 8)  # prime factor 37 > 31
    rc = L.inqamd_hamiltonian_create(ctypes.byref(d), None, ctypes.byref(h))
    assert rc == _lib.ERR_UNSUPPORTED
    assert b"prime factor" in L.inqamd_last_error(None)


def test_null_handle_calls_fail_cleanly():
    L = _lib.lib()
    assert L.inqamd_hamiltonian_apply(None, None, None, 1, 1) == _lib.ERR_ARGUMENT
    assert L.inqamd_set_kpoint(None, None) == _lib.ERR_ARGUMENT
    L.inqamd_hamiltonian_destroy(None)


@pytest.mark.parametrize("bad", ["radius", "capacity"])
def test_sphere_builder_argument_checks(bad):
    L = _lib.lib()
    cnt = ctypes.c_int(0)
    lat = _lib.dvec([5, 0, 0, 0, 5, 0, 0, 0, 5])
    if bad == "radius":
        rc = L.inqamd_sphere_build(lat, _lib.ivec([8, 8, 8]), _lib.dvec([0, 0, 0]), -1.0, 0, None,
                                   None, None, ctypes.byref(cnt))
        assert rc == _lib.ERR_ARGUMENT
    else:
        rc = L.inqamd_sphere_build(lat, _lib.ivec([8, 8, 8]), _lib.dvec([0, 0, 0]), 1.5, 0, None,
                                   None, None, ctypes.byref(cnt))
        assert rc == _lib.ERR_MEMORY and cnt.value > 0
