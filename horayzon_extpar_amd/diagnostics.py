"""Error-analysis helpers around the horizon path (host side).

``distance_to_border`` restates functions.py:58-97 of the reference: the
chord distance [km] from every cell circumcentre to the nearest ICON vertex
on the mesh border (a vertex with a missing adjacent cell, -2 in
cells_of_vertex), on a sphere of radius 6371.0087714 km.  Cells close to the
border can see through the holes of the grid_type-1 mesh there, which is
where the reference walk becomes path dependent.

``border_affected`` marks, per azimuth, the angles whose 40 km ray track can
come near the border: the set outside which the reference walk and the
clip-aware lattice search agree (used to report learned-vs-walk
disagreements).
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
