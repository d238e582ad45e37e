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
    v = np.zeros(1031 * 64)
    d.vloc = v.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    d.max_bands = 4
    d.n = (ctypes.c_int * 3)(1031, 8, 8)  # a prime axis longer than an LDS pencil holds
    rc = L.inqamd_hamiltonian_create(ctypes.byref(d), None, ctypes.byref(h))
    assert rc == _lib.ERR_UNSUPPORTED
    assert b"too long" in L.inqamd_last_error(None)


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


def test_ctypes_structures_match_the_header_layout(tmp_path):
    """sizeof and field offsets of the descriptor structs, from gcc on the header."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    structs = {"inqamd_system_desc": _lib.SystemDesc,
               "inqamd_vector_potential_series": _lib.VectorPotentialSeries,
               "inqamd_td_desc": _lib.TDDesc}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "inq_amd.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    for line in out.splitlines():
        cname, what, val = line.split()
        py = structs[cname]
        expect = ctypes.sizeof(py) if what == "size" else getattr(py, what).offset
        assert int(val) == expect, (cname, what)


def test_new_entry_points_reject_null_handles():
    L = _lib.lib()
    out = ctypes.c_void_p()
    d = _lib.TDDesc()
    assert L.inqamd_td_create(None, ctypes.byref(d), ctypes.byref(out)) == _lib.ERR_ARGUMENT
    assert L.inqamd_td_forward(None, None, None, 1, None, None) == _lib.ERR_ARGUMENT
    assert L.inqamd_td_backward(None, 1.0, None, None, None) == _lib.ERR_ARGUMENT
    assert L.inqamd_td_memory(None) == 0
    L.inqamd_td_destroy(None)
    j = (ctypes.c_double * 3)()
    assert L.inqamd_current(None, None, 1, 1, 2.0, j, j) == _lib.ERR_ARGUMENT
    assert L.inqamd_batch_create(None, 0, None, ctypes.byref(out)) == _lib.ERR_ARGUMENT
    assert L.inqamd_batch_apply(None, None, None, None, None) == _lib.ERR_ARGUMENT
    L.inqamd_batch_destroy(None)
    ms = (ctypes.c_double * 4)()
    n = (ctypes.c_int * 4)()
    assert L.inqamd_hamiltonian_profile(None, None, None, 1, 1, ms, n) == _lib.ERR_ARGUMENT
