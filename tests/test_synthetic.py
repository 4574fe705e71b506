"""Synthetic ICON-like grids and terrains (CPU)."""

import numpy as np

from horayzon_extpar_amd import synthetic


def _check_cov(d):
    cov = d.cells_of_vertex
    assert cov.dtype == np.int32 and cov.shape == (6, d.nvert)
    valid = cov[cov != -2]
    assert valid.min() >= 0 and valid.max() < d.ncell
    # every triangle cell touches exactly three ICON vertices
    assert np.array_equal(np.bincount(valid, minlength=d.ncell), np.full(d.ncell, 3))


def test_config_sizes(c1, c2):
    assert (c1.ncell, c1.nvert) == (5700, 2958)
    assert (c2.ncell, c2.nvert) == (553200, 277662)
    _check_cov(c1)
    assert c2.hsurf.max() == 4500.0 and c2.hsurf.min() == 0.0
    assert c1.hsurf.min() == 0.0 and 0 < c1.hsurf.std() < 1000


def test_interior_vertices_are_six_valent(c1):
    nv = np.sum(c1.cells_of_vertex != -2, axis=0)
    assert np.mean(nv == 6) > 0.9 and nv.max() == 6


def test_deterministic():
    a = synthetic.config1(size_km=30.0, seed=7)
    b = synthetic.config1(size_km=30.0, seed=7)
    c = synthetic.config1(size_km=30.0, seed=8)
    for x, y in zip(a.args(), b.args()):
        assert np.array_equal(x, y)
    assert not np.array_equal(a.hsurf, c.hsurf)


def test_circumcentres_equidistant():
    d = synthetic.lattice_domain(20.0, 20.0, 2.0, shuffle=False)
    u = lambda lon, lat: np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon),  # noqa
                                   np.sin(lat)], -1)
    c = u(d.clon, d.clat)
    v = u(d.vlon, d.vlat)
    cov = d.cells_of_vertex
    for vert in range(0, d.nvert, 7):
        cells = cov[:, vert][cov[:, vert] != -2]
        ang = np.arccos(np.clip(c[cells] @ v[vert], -1, 1))
        assert np.allclose(ang * 6371229.0, 2000.0 / np.sqrt(3), rtol=0.01)


def test_tilted_plane_is_planar():
    d = synthetic.tilted_plane(20.0, 1.0, 0.05, 0.02)
    u = np.stack([np.cos(d.clat) * np.cos(d.clon), np.cos(d.clat) * np.sin(d.clon),
                  np.sin(d.clat)], 1)
    p = (synthetic.RAD_EARTH + d.hsurf)[:, None] * u
    m = np.array(d.meta["plane_normal"])
    r = (p @ m) / np.linalg.norm(m)
    assert np.ptp(r) < 1e-6
