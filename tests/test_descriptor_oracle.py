"""Pins the descriptor restatement (oracle/descriptor.py) with closed forms,
on flat synthetic scenes (up vector +z, north +y, so east is +x):

* the raster of a plane equals the plane at the pixel centres inside the
  mesh and is empty outside it;
* seen from a point on a tilted plane z = a x + b y + c, the bin of azimuth
  phi holds atan(a sin phi + b cos phi): this fixes the azimuth convention
  (clockwise from north) and the sign of the angle;
* from the centre of a bowl z = kappa r^2 / 2 the bin of ring k holds the
  elevation of its farthest sample, atan(kappa R_k / 2), and from the top of
  a dome z = -kappa r^2 / 2 that of its nearest sample: this fixes the ring
  radii R_k = 500 m * 80^(k/7) and the sample spacing;
* bins whose samples all leave the mesh hold -90 deg.

The tolerances are the closed forms' own: a sample moves to the nearest
pixel centre (at most pix / sqrt(2) away), which changes its height by at
most |grad z| pix / sqrt(2), and the BF16 result carries 2^-9 relative.
"""

import math

import numpy as np
import pytest

import descriptor as dsc

F = np.float32


def grid_mesh(w, spacing, zfun):
    n = int(round(w / spacing)) + 1
    x, y = np.meshgrid(np.arange(n) * spacing, np.arange(n) * spacing, indexing="xy")
    x, y = x.ravel(), y.ravel()
    vert = np.stack([x, y, zfun(x, y)], 1).astype(F)
    i = np.arange(n - 1)
    a = (i[None, :] + n * i[:, None]).ravel()
    tri = np.concatenate([np.stack([a, a + 1, a + n + 1], 1), np.stack([a, a + n + 1, a + n], 1)])
    return vert, tri.astype(np.int32)


def flat_frames(pts, zfun):
    org = np.array([[x, y, zfun(np.array(x), np.array(y))] for x, y in pts], F)
    n = np.tile([0.0, 0.0, 1.0], (len(pts), 1))
    north = np.tile([0.0, 1.0, 0.0], (len(pts), 1))
    return org, n, north


def values(desc):
    return dsc.bf16_to_f32(desc).astype(np.float64).reshape(-1, 24, 8) * 10.0  # degrees


def test_tables():
    dist, sn, cs = dsc.tables()
    # three samples per ring, the last on the ring's outer radius
    outer = np.array([500.0 * 80.0 ** (k / 7.0) for k in range(8)])
    assert np.allclose(dist[2::3], outer, rtol=1e-7) and abs(dist[-1] - 40000.0) < 1e-2
    assert np.all(np.diff(dist) > 0) and abs(dist[0] - 250.0 * 2.0 ** (1.0 / 3.0)) < 1e-3
    assert abs(sn[6] - 1.0) < 1e-7 and abs(cs[0] - 1.0) < 1e-7 and abs(cs[12] + 1.0) < 1e-7


@pytest.fixture(scope="module")
def plane():
    a, b, c = 0.05, -0.03, 500.0
    zf = lambda x, y: a * x + b * y + c  # noqa: E731
    vert, tri = grid_mesh(30000.0, 100.0, zf)
    return (a, b, c), zf, vert, dsc.raster(vert, tri)


def test_raster_of_plane(plane):
    (a, b, c), zf, vert, ras = plane
    z, pix = ras["z"], float(ras["pix"])
    assert 60.0 < pix < 80.0  # sqrt(area / ntri) for 100 m cells
    ny, nx = z.shape
    xc = float(ras["x0"]) + np.arange(nx) * pix
    yc = float(ras["y0"]) + np.arange(ny) * pix
    X, Y = np.meshgrid(xc, yc)
    inside = (X < 30000.0 - 1e-3 * pix) & (Y < 30000.0 - 1e-3 * pix)
    assert np.all(np.isfinite(z[inside])) and not np.any(np.isfinite(z[~inside]))
    assert np.abs(z[inside] - zf(X[inside], Y[inside])).max() < 1e-2


def test_tilted_plane_slope_per_azimuth(plane):
    (a, b, c), zf, vert, ras = plane
    pts = [(15000.0, 15000.0), (14321.5, 15877.25), (16100.0, 13950.0)]
    org, n, north = flat_frames(pts, zf)
    v = values(dsc.descriptors(ras, org, n, north))
    phi = np.arange(24) * (2.0 * math.pi / 24)
    s = a * np.sin(phi) + b * np.cos(phi)
    want = np.degrees(np.arctan(s))
    dist = dsc.tables()[0].astype(np.float64)
    g = math.hypot(a, b)
    pix = float(ras["pix"])
    for k in range(6):  # rings up to 11.4 km stay on the 30 km plane
        eps = np.degrees(g * pix / math.sqrt(2.0) / dist[3 * k])
        tol = eps + np.abs(want) * 2.0 ** -8 + 1e-4
        err = np.abs(v[:, :, k] - want[None, :])
        assert np.all(err <= tol[None, :]), (k, err.max(), tol.max())
    # swapped sine / cosine (azimuth from east) would miss by up to ~4 deg
    wrong = np.degrees(np.arctan(a * np.cos(phi) + b * np.sin(phi)))
    assert np.abs(v[:, :, 5] - wrong[None, :]).max() > 1.0


@pytest.fixture(scope="module", params=[1.0, -1.0], ids=["bowl", "dome"])
def paraboloid(request):
    sign, kappa, cx = request.param, 1e-4, 12000.0
    zf = lambda x, y: sign * kappa * ((x - cx) ** 2 + (y - cx) ** 2) / 2.0  # noqa: E731
    vert, tri = grid_mesh(24000.0, 50.0, zf)
    return sign, kappa, cx, zf, dsc.raster(vert, tri)


def test_paraboloid_ring_radii(paraboloid):
    sign, kappa, cx, zf, ras = paraboloid
    org, n, north = flat_frames([(cx, cx)], zf)
    v = values(dsc.descriptors(ras, org, n, north))[0]
    dist = dsc.tables()[0].astype(np.float64)
    pix = float(ras["pix"])
    dl = pix / math.sqrt(2.0)
    for k in range(6):
        d = dist[3 * k + 2] if sign > 0 else dist[3 * k]  # farthest / nearest sample
        want = math.degrees(math.atan(sign * kappa * d / 2.0))
        # height error of a sample moved by dl: kappa (d dl + dl^2 / 2); per metre of distance
        tol = math.degrees(kappa * (d * dl + dl * dl / 2.0) / d) + abs(want) * 2.0 ** -8 + 1e-4
        assert np.all(np.abs(v[:, k] - want) <= tol), (k, v[:, k], want, tol)
        # from ring 1 on, the ring's middle sample is more than two
        # tolerances away: the bins then single out the right sample
        other = math.degrees(math.atan(sign * kappa * dist[3 * k + 1] / 2.0))
        if k >= 1:
            assert abs(other - want) > 2.0 * tol
            assert np.all(np.abs(v[:, k] - other) > tol)
    # ring 7 (from 26 km) leaves the 24 km mesh in every direction, ring 6
    # (from 14 km) along the axes; along the diagonals (17 km) ring 6 stays on it
    m90 = float(dsc.bf16_to_f32(dsc.bf16_bits(np.full(1, F(-90.0) * F(0.1))))[0]) * 10.0
    assert np.all(v[:, 7] == m90) and np.all(v[0::6, 6] == m90)
    assert np.all(v[3::6, 6] > -90.0)


def test_empty_scene_gives_minus_ninety():
    vert = np.zeros((0, 3), F)
    tri = np.zeros((0, 3), np.int32)
    ras = dsc.raster(vert, tri)
    org, n, north = flat_frames([(0.0, 0.0)], lambda x, y: 0.0 * x)
    d = dsc.descriptors(ras, org, n, north)
    assert np.all(d == dsc.bf16_bits(np.full(1, F(-90.0) * F(0.1)))[0])
