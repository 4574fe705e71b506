"""Host diagnostics: the observer-perspective transform, the sectional
terrain horizon built on it, and the per-azimuth border flags.

* observer_perspective (functions.py:11-54) against closed forms on the
  sphere: a point straight above the observer, and points on the equator
  and on a meridian at the observer's elevation, whose elevation angle is
  minus half their central angle and whose distance is the chord;
* section_horizon from the observer coordinates of the C1 mesh equals the
  oracle's FP64 half-plane horizon (orc_true_horizon, which works on the FP32
  ENU scene) to the FP32 rounding of that scene;
* border_azimuths on a hand-built two-vertex case.
"""

import numpy as np
import pytest

import oracle
from horayzon_extpar_amd import diagnostics as dg
from horayzon_extpar_amd import synthetic
from horayzon_extpar_amd.horizon_svf import scene_export

R = 6371229.0


def test_observer_perspective_closed_forms():
    lon0, lat0, h0 = np.radians(8.3), np.radians(46.7), 1500.0
    phi, theta, r = dg.observer_perspective(np.array([lon0]), np.array([lat0]),
                                            np.array([h0 + 100.0]), lon0, lat0, h0)
    assert abs(theta[0] - 90.0) < 1e-9 and abs(r[0] - 100.0) < 1e-6
    g = np.array([1e-4, 1e-3, 5e-3])  # central angles [rad]
    h = 300.0
    # on the equator, east and west of the observer
    for sgn, az in ((1.0, 90.0), (-1.0, 270.0)):
        phi, theta, r = dg.observer_perspective(sgn * g, np.zeros(3), np.full(3, h), 0.0, 0.0, h)
        assert np.allclose(phi, az, atol=1e-9)
        assert np.allclose(theta, -np.degrees(g / 2), rtol=0, atol=1e-9)
        assert np.allclose(r, 2 * (R + h) * np.sin(g / 2), rtol=1e-12)
    # on the observer's meridian, north and south
    for sgn, az in ((1.0, 0.0), (-1.0, 180.0)):
        phi, theta, r = dg.observer_perspective(np.full(3, lon0), lat0 + sgn * g, np.full(3, h),
                                                lon0, lat0, h)
        d = np.minimum(np.abs(phi - az), 360.0 - np.abs(phi - az))
        assert np.all(d < 1e-9)
        assert np.allclose(theta, -np.degrees(g / 2), rtol=0, atol=1e-9)
        assert np.allclose(r, 2 * (R + h) * np.sin(g / 2), rtol=1e-12)


def test_mesh_vertices_match_scene(c1):
    lon, lat, hgt = dg.mesh_vertices(*c1.args(), grid_type=1)
    sc = scene_export(*c1.args(), grid_type=1)
    assert lon.size == sc["vert32"].shape[0]
    assert sc["tri"].max() < lon.size
    # the scene's ENU vertices, seen from its ENU origin, have the same
    # distances as the mesh vertices seen from the same point
    lon0, lat0 = sc["lon0_lat0"]
    _, _, r = dg.observer_perspective(lon, lat, hgt, lon0, lat0, 0.0)
    rs = np.linalg.norm(sc["vert32"].astype(np.float64), axis=1)
    assert np.max(np.abs(r - rs)) < 0.05  # FP32 rounding of coordinates up to ~70 km


@pytest.mark.parametrize("num_hori", [24, 48])
def test_section_horizon_equals_oracle_true_horizon(c1, num_hori):
    lon, lat, hgt = dg.mesh_vertices(*c1.args(), grid_type=1)
    tri = scene_export(*c1.args(), grid_type=1)["tri"]
    cells = np.arange(7, c1.ncell, 611)
    th = np.degrees(oracle.true_horizon(c1, cells, num_hori=num_hori, grid_type=1))
    az = np.arange(num_hori) * (360.0 / num_hori)
    worst = 0.0
    for q, i in enumerate(cells):
        phi, theta, r = dg.observer_perspective(lon, lat, hgt, c1.clon[i], c1.clat[i],
                                                c1.hsurf[i] + 0.2)
        sh = dg.section_horizon(phi, theta, r, tri, az, 40000.0)
        void = th[q] <= -89.999
        assert np.array_equal(void, sh <= -89.999)
        worst = max(worst, float(np.abs(sh[~void] - th[q][~void]).max(initial=0.0)))
    # the oracle works on the FP32 scene (mm on vertices km away): measured
    # 6.5e-6 deg at most on these cells
    assert worst < 5e-5, worst


def test_triangles_below_the_horizontal_never_raise_it(c1):
    # a triangle whose corners all lie below the observer's horizontal plane
    # lies below it everywhere: dropping those (the theta part of the
    # reference's plot mask) leaves every positive horizon angle unchanged
    lon, lat, hgt = dg.mesh_vertices(*c1.args(), grid_type=1)
    tri = scene_export(*c1.args(), grid_type=1)["tri"]
    i = c1.ncell // 2
    phi, theta, r = dg.observer_perspective(lon, lat, hgt, c1.clon[i], c1.clat[i],
                                            c1.hsurf[i] + 0.2)
    az = np.arange(24) * 15.0
    full = dg.section_horizon(phi, theta, r, tri, az, 40000.0)
    up = theta[tri].max(axis=1) > 0.0
    assert 0 < up.sum() < len(tri)
    part = dg.section_horizon(phi, theta, r, tri[up], az, 40000.0)
    above = full > 0.0
    assert above.any() and np.array_equal(full[above], part[above])


def test_triangle_mask_parts():
    # corners: 0-2 a triangle north-east above the horizontal; 3-5 one that
    # straddles north (azimuths 359 / 1 deg); 6-8 one beyond 40 km; 9-11 one
    # below the horizontal
    phi = np.array([30, 40, 35, 359, 1, 0.5, 30, 40, 35, 30, 40, 35], float)
    theta = np.array([1, 2, 1.5, 1, 1, 2, 1, 2, 1.5, -1, -2, -0.5], float)
    r = np.array([5e3, 6e3, 7e3, 5e3, 5e3, 6e3, 41e3, 42e3, 45e3, 5e3, 6e3, 7e3], float)
    tri = np.arange(12).reshape(4, 3)
    assert dg.triangle_mask(phi, theta, r, tri, 40000.0).tolist() == [True, False, False, False]


def test_border_azimuths_two_vertices():
    # one cell at (0, 0); a border vertex 10 km due north, an interior
    # vertex 10 km due east
    d = 10000.0 / R
    clon, clat = np.array([0.0]), np.array([0.0])
    vlon, vlat = np.array([0.0, d]), np.array([d, 0.0])
    cov = np.array([[0, 0]] * 5 + [[-2, 0]], np.int32)
    az = np.arange(24) * 15.0
    f = dg.border_azimuths(clon, clat, vlon, vlat, cov, [0], az, 40000.0, 1000.0)[0]
    # within 1 km of the segment: |sin(az)| * 10 km < 1 km and pointing north
    exp = (np.abs(np.sin(np.radians(az))) * 10000.0 < 1000.0) & (np.cos(np.radians(az)) > 0)
    assert np.array_equal(f, exp)
    f = dg.border_azimuths(clon, clat, vlon, vlat, cov, [0], az, 5000.0, 1000.0)[0]
    assert not f.any()  # the segment ends 5 km short of it
    f = dg.border_azimuths(clon, clat, vlon, vlat, cov, [0], az, 40000.0, 10500.0)[0]
    assert f.all()  # the observer itself is within the margin
