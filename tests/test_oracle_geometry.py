"""Pins the sphere/projector geometry (oracle and the library's host builders).

Reference known answers: sphere point counts and volume ratios
(src/basis/spherical_grid.hpp:272-396) and containing_cube bounds
(src/basis/containing_cube.hpp:77-117).  The library's builders must agree
with the oracle bit for bit (FP32 geometry, point order).  The projector
table is checked against its closed form f(d) Y_lm(r).
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from inq_amd import geometry
from helpers import grid_positions

HERE = os.path.dirname(os.path.abspath(__file__))
KNOWN = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))


def _grid(lat, spacing):
    return [int(round(np.linalg.norm(lat[i]) / spacing)) for i in range(3)]


@pytest.mark.parametrize("builder", ["oracle", "library"])
def test_sphere_counts_cubic(builder):
    k = KNOWN["sphere"]["cubic"]
    lat = np.eye(3) * k["cell"]
    n = _grid(lat, k["spacing"])
    vol = k["cell"] ** 3
    for c in k["centres"]:
        if builder == "oracle":
            coords, dist, rel = O.sphere_build(lat, n, c, k["radius"])
        else:
            coords, dist, rel = geometry.spherical_grid(lat, n, c, k["radius"])
        assert dist.size == k["count"]
        ratio = dist.size * vol / np.prod(n) / (4.0 / 3.0 * np.pi * k["radius"] ** 3)
        assert ratio == pytest.approx(k["volume_ratio"], abs=5e-11)


@pytest.mark.parametrize("builder", ["oracle", "library"])
def test_sphere_count_noncubic(builder):
    k = KNOWN["sphere"]["noncubic"]
    lat = np.array(k["lattice"])
    n = _grid(lat, k["spacing"])
    f = O.sphere_build if builder == "oracle" else geometry.spherical_grid
    coords, dist, rel = f(lat, n, k["centre"], k["radius"])
    assert dist.size == k["count"]
    vol = abs(np.linalg.det(lat))
    ratio = dist.size * vol / np.prod(n) / (4.0 / 3.0 * np.pi * k["radius"] ** 3)
    assert ratio == pytest.approx(k["volume_ratio"], abs=5e-11)


@pytest.mark.parametrize("builder", ["oracle", "library"])
def test_containing_cube(builder):
    f = O.containing_cube if builder == "oracle" else geometry.containing_cube
    o = KNOWN["containing_cube"]["orthogonal"]
    lat = np.diag(o["lengths"]).astype(float)
    lo, hi = f(lat, _grid(lat, o["spacing"]), o["centre"], o["radius"])
    assert lo == o["lo"] and hi == o["hi"]
    c = KNOWN["containing_cube"]["fcc"]
    aa = c["aa"]
    lat = np.array([[0, aa / 2, aa / 2], [aa / 2, 0, aa / 2], [aa / 2, aa / 2, 0]])
    lo, hi = f(lat, _grid(lat, c["spacing"]), c["centre"], c["radius"])
    assert lo == c["lo"] and hi == c["hi"]


@pytest.mark.parametrize("centre", [(0.0, 0.0, 0.0), (2.565, 2.565, 2.565), (-4.9, 1.3, 5.13),
                                    (10.1, -0.2, 3.3)])
def test_library_builder_matches_oracle_bitwise(centre):
    lat = np.array([[10.26, 0, 0], [0.3, 9.9, 0], [0.1, -0.2, 10.5]])
    n = (32, 30, 33)
    a = O.sphere_build(lat, n, centre, 1.8)
    b = geometry.spherical_grid(lat, n, centre, 1.8)
    for x, y in zip(a, b):
        assert x.shape == y.shape and (x == y).all()
    ma = O.gaussian_projector_matrix(lat, a[1], a[2], 0.9, 4)
    mb = geometry.gaussian_projector_matrix(lat, b[1], b[2], 0.9, 4)
    assert (ma == mb).all()


def test_sphere_points_are_the_ball():
    """Every grid point within R of the centre's periodic images, no others."""
    lat = np.eye(3) * 6.0
    n = (16, 16, 16)
    centre = np.array([2.9, -2.95, 0.1])
    R = 1.9
    coords, dist, rel = geometry.spherical_grid(lat, n, centre, R)
    r = grid_positions(lat, n)
    d = r - centre
    d -= 6.0 * np.round(d / 6.0)
    inside = set(map(tuple, np.argwhere((d * d).sum(-1) <= R * R)))
    assert set(map(tuple, coords)) == inside
    # relative positions are r - R_image in contravariant (fractional) units
    cart = rel.astype(float) @ lat
    assert np.allclose(np.linalg.norm(cart, axis=1), dist, atol=1e-5)


def test_projector_matrix_closed_form():
    lat = np.eye(3) * 8.0
    n = (20, 20, 20)
    coords, dist, rel = geometry.spherical_grid(lat, n, (0.37, -0.2, 0.11), 2.0)
    m = geometry.gaussian_projector_matrix(lat, dist, rel, 0.9, 4)
    x = rel.astype(float) @ lat
    r = np.linalg.norm(x, axis=1)
    f = np.exp(-dist.astype(float) ** 2 / 0.81)
    assert np.allclose(m[0], f / np.sqrt(4 * np.pi), rtol=1e-14)
    y1 = np.sqrt(3 / (4 * np.pi))
    for row, comp in ((1, 1), (2, 2), (3, 0)):
        assert np.allclose(m[row], f * y1 * x[:, comp] / r, rtol=1e-13, atol=1e-15)
