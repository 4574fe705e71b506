"""Error-analysis helpers around the horizon path (host side).

``distance_to_border`` restates functions.py:58-97 of the reference: the
chord distance [km] from every cell circumcentre to the nearest ICON vertex
on the mesh border (a vertex with a missing adjacent cell, -2 in
cells_of_vertex), on a sphere of radius 6371.0087714 km.  Cells close to the
border can see through the holes of the grid_type-1 mesh there, which is
where the reference walk becomes path dependent.

``observer_perspective`` restates functions.py:11-54: points given by
longitude, latitude and elevation, seen from an observer, as azimuth
(clockwise from north), elevation angle and distance, through spherical
ECEF (R = 6371229 m) and the observer's topocentric east-north-up frame.
The reference calls pyproj's cart + topocentric pipeline on a sphere; the
same three steps are written out here (pyproj is not a dependency).
``mesh_vertices`` gives the longitude, latitude and elevation of the
vertices the grid_type 0 / 1 triangle mesh is built from (the reference's
clon_ext / clat_ext / hsurf_ext, mo_lradtopo_horayzon.cpp:240-300), and
``triangle_mask`` selects the triangles the reference's verification plots
draw in that frame (verification.py:321-333).  ``section_horizon`` takes the
highest point of the mesh in each azimuth's vertical half-plane within the
search distance, from those observer coordinates: the FP64 terrain horizon
that the walk's 0.5 deg lattice value approximates.

``border_azimuths`` marks, per cell and azimuth, the search segments that
pass within a margin of a mesh-boundary vertex: the angles where the
reference walk can see through a hole of the mesh and its result depends on
the path it took, reported apart from the other angles when the walk is
compared with the FP64 terrain horizon.
"""

from __future__ import annotations

import numpy as np

RAD_E_KM = 6371.0087714  # functions.py:84


def _xyz(lon, lat, r):
    return np.stack([r * np.cos(lat) * np.cos(lon), r * np.cos(lat) * np.sin(lon),
                     r * np.sin(lat)], axis=1)


def distance_to_border(clon, clat, vlon, vlat, cells_of_vertex) -> np.ndarray:
    """Chord distance [km] of each cell circumcentre to the nearest border
    vertex (functions.py:58-97); inf when the mesh has no border."""
    from scipy.spatial import cKDTree

    cov = np.asarray(cells_of_vertex)
    border = np.any(cov == -2, axis=0)
    q = _xyz(np.asarray(clon, np.float64), np.asarray(clat, np.float64), RAD_E_KM)
    if not border.any():
        return np.full(q.shape[0], np.inf)
    pts = _xyz(np.asarray(vlon, np.float64)[border], np.asarray(vlat, np.float64)[border],
               RAD_E_KM)
    dist, _ = cKDTree(pts).query(q, k=1)
    return dist


R_EARTH_M = 6371229.0  # functions.py:41-42 (proj sphere radius)


def _ecef(lon, lat, h):
    lon, lat, h = (np.asarray(x, np.float64) for x in (lon, lat, h))
    r = R_EARTH_M + h
    return (r * np.cos(lat) * np.cos(lon), r * np.cos(lat) * np.sin(lon), r * np.sin(lat))


def observer_perspective(lon, lat, elevation, lon_obs, lat_obs, elevation_obs):
    """Points as seen from an observer (functions.py:11-54).

    lon, lat [rad], elevation [m] of the points; the observer at lon_obs,
    lat_obs [rad], elevation_obs [m].  Returns (phi, theta, radius): azimuth
    [deg] clockwise from north in [0, 360), elevation angle [deg] in
    [-90, 90], distance [m]; all on a sphere of radius 6371229 m."""
    x, y, z = _ecef(lon, lat, elevation)
    x0, y0, z0 = _ecef(lon_obs, lat_obs, elevation_obs)
    dx, dy, dz = x - x0, y - y0, z - z0
    so, co = np.sin(lon_obs), np.cos(lon_obs)
    sa, ca = np.sin(lat_obs), np.cos(lat_obs)
    e = -so * dx + co * dy
    n = -sa * co * dx - sa * so * dy + ca * dz
    u = ca * co * dx + ca * so * dy + sa * dz
    radius = np.sqrt(e * e + n * n + u * u)
    with np.errstate(invalid="ignore", divide="ignore"):
        theta = 90.0 - np.degrees(np.arccos(u / radius))
    phi = np.degrees(np.arctan2(e, n))
    phi = np.where(phi < 0.0, phi + 360.0, phi)
    return phi, theta, radius


def mesh_vertices(clon, clat, hsurf, vlon, vlat, cells_of_vertex, grid_type=1):
    """(lon, lat, elevation) of the mesh vertices, in the order the
    triangle indices of hz_scene_export refer to: the cells, then (grid_type
    1) every ICON vertex with six cells, at the mean elevation of its cells
    (mo_lradtopo_horayzon.cpp:280-300)."""
    clon, clat, hsurf = (np.asarray(x, np.float64) for x in (clon, clat, hsurf))
    if grid_type == 0:
        return clon, clat, hsurf
    cov = np.asarray(cells_of_vertex)
    six = np.flatnonzero(np.all(cov != -2, axis=0))
    hs = np.zeros(six.size)
    for j in range(6):  # the reference's summation order (cells_of_vertex rows)
        hs = hs + hsurf[cov[j, six]]
    return (np.concatenate([clon, np.asarray(vlon, np.float64)[six]]),
            np.concatenate([clat, np.asarray(vlat, np.float64)[six]]),
            np.concatenate([hsurf, hs / 6.0]))


def triangle_mask(phi, theta, radius, tri, dist_search):
    """Triangles (rows of tri, vertex indices) that the reference's
    verification draws in the observer frame (verification.py:329-333): a
    corner within dist_search, a corner above the horizontal, and an azimuth
    spread below 180 deg (the triangle does not wrap around the observer)."""
    tri = np.asarray(tri)
    return ((radius[tri].min(axis=1) <= dist_search) & (theta[tri].max(axis=1) > 0.0)
            & (np.ptp(phi[tri], axis=1) < 180.0))


def section_horizon(phi, theta, radius, tri, azimuth_deg, dist_search):
    """Highest elevation angle [deg] of the mesh within dist_search in the
    vertical half-plane of each azimuth (deg clockwise from north), from the
    observer coordinates of its vertices; -90 where the half-plane meets no
    triangle.  Each triangle is cut by the plane (an edge crossing it, or a
    corner on it), the cut is kept on the azimuth's side of the observer and
    clipped to the sphere of radius dist_search about the observer; the
    highest of the cut's end points is the triangle's contribution."""
    tri = np.asarray(tri)
    t, p = np.radians(theta), np.radians(phi)
    rh = radius * np.cos(t)
    E, N, U = rh * np.sin(p), rh * np.cos(p), radius * np.sin(t)
    R2 = float(dist_search) ** 2
    out = np.full(len(azimuth_deg), -90.0)
    for j, az in enumerate(np.radians(np.asarray(azimuth_deg, np.float64))):
        hx, hy = np.sin(az), np.cos(az)
        u_all = E * hx + N * hy          # along the azimuth
        w_all = E * hy - N * hx          # across the plane
        uu, ww, zz = u_all[tri], w_all[tri], U[tri]  # [ntri, 3]
        pts_u, pts_z, valid = [], [], []
        for k in range(3):
            a, b = k, (k + 1) % 3
            on = ww[:, a] == 0.0
            pts_u.append(uu[:, a])
            pts_z.append(zz[:, a])
            valid.append(on)
            cross = ((ww[:, a] < 0) & (ww[:, b] > 0)) | ((ww[:, a] > 0) & (ww[:, b] < 0))
            with np.errstate(invalid="ignore", divide="ignore"):
                s = ww[:, a] / (ww[:, a] - ww[:, b])
            pts_u.append(uu[:, a] + s * (uu[:, b] - uu[:, a]))
            pts_z.append(zz[:, a] + s * (zz[:, b] - zz[:, a]))
            valid.append(cross)
        pu, pz, pv = np.stack(pts_u, 1), np.stack(pts_z, 1), np.stack(valid, 1)
        sel = np.flatnonzero(pv.sum(1) >= 2)
        if sel.size == 0:
            continue
        # the first two valid points of each cut
        order = np.argsort(~pv[sel], axis=1, kind="stable")[:, :2]
        rows = sel[:, None]
        a0, a1 = pu[rows, order[:, :1]][:, 0], pu[rows, order[:, 1:]][:, 0]
        z0, z1 = pz[rows, order[:, :1]][:, 0], pz[rows, order[:, 1:]][:, 0]
        keep = ~((a0 < 0) & (a1 < 0))
        a0, a1, z0, z1 = a0[keep], a1[keep], z0[keep], z1[keep]
        # keep the part with u >= 0
        with np.errstate(invalid="ignore", divide="ignore"):
            s = a0 / (a0 - a1)
        zc = z0 + s * (z1 - z0)
        m0, m1 = a0 < 0, a1 < 0
        a0, z0 = np.where(m0, 0.0, a0), np.where(m0, zc, z0)
        a1, z1 = np.where(m1, 0.0, a1), np.where(m1, zc, z1)
        # clip to |(u, z)| <= dist_search
        du, dz = a1 - a0, z1 - z0
        A = du * du + dz * dz
        B = a0 * du + z0 * dz
        C = a0 * a0 + z0 * z0 - R2
        disc = B * B - A * C
        ok = (A > 0) & (disc >= 0)
        with np.errstate(invalid="ignore", divide="ignore"):
            sq = np.sqrt(np.maximum(disc, 0.0))
            s_lo = np.clip((-B - sq) / A, 0.0, 1.0)
            s_hi = np.clip((-B + sq) / A, 0.0, 1.0)
        ok &= s_lo < s_hi
        inside0 = a0 * a0 + z0 * z0 <= R2
        point = (A == 0) & inside0  # a cut of one point
        e_best = np.full(a0.size, -np.inf)
        for sv in (s_lo, s_hi):
            uu_ = a0 + sv * du
            zz_ = z0 + sv * dz
            e_best = np.where(ok, np.maximum(e_best, np.arctan2(zz_, uu_)), e_best)
        e_best = np.where(point, np.maximum(e_best, np.arctan2(z0, a0)), e_best)
        if np.isfinite(e_best).any():
            out[j] = max(out[j], np.degrees(e_best[np.isfinite(e_best)].max()))
    return out


def border_azimuths(clon, clat, vlon, vlat, cells_of_vertex, cells, azimuth_deg,
                    dist_search, margin_m):
    """[len(cells), len(azimuth_deg)] flags: the horizontal search segment
    of the azimuth (length dist_search from the cell circumcentre) passes
    within margin_m of a mesh-boundary vertex (a vertex with fewer than six
    cells, -2 in cells_of_vertex), where the grid_type-1 mesh has a hole."""
    cov = np.asarray(cells_of_vertex)
    bnd = np.flatnonzero(np.any(cov == -2, axis=0))
    cells = np.asarray(cells)
    az = np.radians(np.asarray(azimuth_deg, np.float64))
    out = np.zeros((cells.size, az.size), bool)
    if bnd.size == 0:
        return out
    blon, blat = np.asarray(vlon, np.float64)[bnd], np.asarray(vlat, np.float64)[bnd]
    zero = np.zeros(bnd.size)
    for q, i in enumerate(cells):
        phi, theta, r = observer_perspective(blon, blat, zero, clon[i], clat[i], 0.0)
        rh = r * np.cos(np.radians(theta))
        near = rh <= dist_search + margin_m
        if not near.any():
            continue
        p, rh = np.radians(phi[near]), rh[near]
        d = p[None, :] - az[:, None]
        along = rh[None, :] * np.cos(d)
        across = np.abs(rh[None, :] * np.sin(d))
        dist = np.where(along < 0, rh[None, :],
                        np.where(along > dist_search,
                                 np.hypot(along - dist_search, across), across))
        out[q] = (dist < margin_m).any(axis=1)
    return out
