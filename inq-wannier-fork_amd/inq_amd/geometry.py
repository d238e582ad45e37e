"""Host-side projector geometry through the library's C ABI (no GPU needed).

Mirrors basis::containing_cube, basis::spherical_grid and projector::build
(src/basis/containing_cube.hpp:20-45, src/basis/spherical_grid.hpp:58-158,
src/hamiltonian/projector.hpp:30-70).
"""
import ctypes

import numpy as np

from . import _lib


def containing_cube(lattice, n, pos, radius):
    lo = (ctypes.c_int * 3)()
    hi = (ctypes.c_int * 3)()
    _lib.check(_lib.lib().inqamd_containing_cube(_lib.dvec(np.ravel(lattice)), _lib.ivec(n),
                                                 _lib.dvec(pos), float(radius), lo, hi))
    return list(lo), list(hi)


def spherical_grid(lattice, n, center, radius):
    """Points of the grid within `radius` of `center` or its periodic images.

    Returns (coords int32 [K,3], distance float32 [K], relpos float32 [K,3]) in
    the reference's point order (replica, then x, y, z of the containing cube).
    """
    L = _lib.lib()
    cnt = ctypes.c_int(0)
    cap = 1024
    while True:
        coords = np.zeros((cap, 3), np.int32)
        dist = np.zeros(cap, np.float32)
        rel = np.zeros((cap, 3), np.float32)
        rc = L.inqamd_sphere_build(_lib.dvec(np.ravel(lattice)), _lib.ivec(n), _lib.dvec(center),
                                   float(radius), cap, coords.ctypes.data, dist.ctypes.data,
                                   rel.ctypes.data, ctypes.byref(cnt))
        if rc == _lib.ERR_MEMORY:
            cap = cnt.value
            continue
        _lib.check(rc)
        k = cnt.value
        return coords[:k].copy(), dist[:k].copy(), rel[:k].copy()


def gaussian_projector_matrix(lattice, distance, relpos, sigma, nlm=4):
    k = int(distance.size)
    m = np.zeros((nlm, max(k, 1)), np.float64)
    d = np.ascontiguousarray(distance, np.float32)
    r = np.ascontiguousarray(relpos, np.float32)
    _lib.check(_lib.lib().inqamd_gaussian_projector_matrix(
        _lib.dvec(np.ravel(lattice)), k, d.ctypes.data, r.ctypes.data, float(sigma), nlm,
        m.ctypes.data, max(k, 1)))
    return m[:, :k].copy()
