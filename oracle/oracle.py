"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only to check or time the HIP library against it.  The
functions follow mo_lradtopo_horayzon.cpp as cited in horayzon_oracle.cpp.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_lib = None


class OracleStats(ctypes.Structure):
    _fields_ = [("num_rays", ctypes.c_longlong), ("num_tri_tests", ctypes.c_longlong),
                ("num_triangles", ctypes.c_longlong),
                ("num_mesh_vertices", ctypes.c_longlong), ("lon0", ctypes.c_double),
                ("lat0", ctypes.c_double), ("t_setup", ctypes.c_double),
                ("t_trace", ctypes.c_double), ("nthreads", ctypes.c_int),
                ("cells_done", ctypes.c_int), ("cells_walked", ctypes.c_int),
                ("pad", ctypes.c_int)]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++, no GPU needed)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


_D = ctypes.c_double
_I = ctypes.c_int
_F = ctypes.c_float


def _inputs(dom):
    clon = np.ascontiguousarray(dom.clon, dtype=np.float64)
    clat = np.ascontiguousarray(dom.clat, dtype=np.float64)
    hsurf = np.ascontiguousarray(dom.hsurf, dtype=np.float64)
    vlon = np.ascontiguousarray(dom.vlon, dtype=np.float64)
    vlat = np.ascontiguousarray(dom.vlat, dtype=np.float64)
    cov = np.ascontiguousarray(dom.cells_of_vertex, dtype=np.int32)
    return clon, clat, hsurf, vlon, vlat, cov


def mesh(dom, grid_type: int):
    """Mesh vertices (lon, lat, elev) and triangle indices."""
    clon, clat, hsurf, vlon, vlat, cov = _inputs(dom)
    nt, nv = _I(), _I()
    L = lib()
    L.orc_mesh(_p(clon, _D), _p(clat, _D), _p(hsurf, _D), _p(vlon, _D), _p(vlat, _D),
               _p(cov, _I), clon.size, vlon.size, grid_type, ctypes.byref(nt),
               ctypes.byref(nv), None, None)
    vert = np.empty((nv.value, 3), np.float64)
    tri = np.empty((nt.value, 3), np.int32)
    L.orc_mesh(_p(clon, _D), _p(clat, _D), _p(hsurf, _D), _p(vlon, _D), _p(vlat, _D),
               _p(cov, _I), clon.size, vlon.size, grid_type, ctypes.byref(nt),
               ctypes.byref(nv), _p(vert, _D), _p(tri, _I))
    return vert, tri


def enu(dom, grid_type: int, ray_org_elev: float = 0.2):
    """ENU prologue: dict with vert_enu (f64), vert32, centre, pole, lon0,
    lat0, frame_n, frame_north (f64) and org (f32) per cell."""
    clon, clat, hsurf, vlon, vlat, cov = _inputs(dom)
    vert, _ = mesh(dom, grid_type)
    nv = vert.shape[0]
    ncell = clon.size
    ve = np.empty((nv, 3), np.float64)
    v32 = np.empty((nv, 3), np.float32)
    cpo = np.empty(8, np.float64)
    fn = np.empty((ncell, 3), np.float64)
    fN = np.empty((ncell, 3), np.float64)
    org = np.empty((ncell, 3), np.float32)
    lib().orc_enu(_p(clon, _D), _p(clat, _D), _p(hsurf, _D), _p(vlon, _D), _p(vlat, _D),
                  _p(cov, _I), ncell, vlon.size, grid_type, _D(ray_org_elev), _p(ve, _D),
                  _p(v32, _F), _p(cpo, _D), _p(fn, _D), _p(fN, _D), _p(org, _F))
    return dict(vert_enu=ve, vert32=v32, centre=cpo[0:3], pole=cpo[3:6], lon0=cpo[6],
                lat0=cpo[7], frame_n=fn, frame_north=fN, org=org)


def horizon_svf(dom, num_hori=24, grid_type=1, dist_search=40000.0, ray_org_elev=0.2,
                refine_factor=1, svf_type=2, mode=0, brute=False, nthreads=0,
                cell_begin=0, cell_end=0, time_budget_s=0.0, out=None, pred=None):
    """Oracle horizon [num_hori, ncell] (deg) and skyview [ncell].
    mode 0 = reference walk, 1 = clip-aware lattice climb, 2 = lattice search
    started at pred [ncell, 24] (deg), 3 / 4 = modes 1 / 2 certified (the
    library's certify=1): the reference walk on cells with a mesh-boundary
    point in reach, the lattice search on the others."""
    clon, clat, hsurf, vlon, vlat, cov = _inputs(dom)
    ncell = clon.size
    if out is None:
        horizon = np.full((num_hori, ncell), np.nan)
        skyview = np.full(ncell, np.nan)
    else:
        horizon, skyview = out
    st = OracleStats()
    pp = None
    if pred is not None:
        pred = np.ascontiguousarray(pred, np.float32)
        assert pred.shape == (ncell, 24)
        pp = _p(pred, _F)
    rc = lib().orc_horizon_svf(_p(clon, _D), _p(clat, _D), _p(hsurf, _D), _p(vlon, _D),
                               _p(vlat, _D), _p(cov, _I), _p(horizon, _D), _p(skyview, _D),
                               ncell, vlon.size, num_hori, grid_type, _D(dist_search),
                               _D(ray_org_elev), refine_factor, svf_type, mode, int(brute),
                               nthreads, cell_begin, cell_end, _D(time_budget_s), pp,
                               ctypes.byref(st))
    if rc != 0:
        raise ValueError("oracle: bad arguments")
    return horizon, skyview, st


def occluded32(vert32, tri, org, dirs, tfar, use_bvh=False):
    v = np.ascontiguousarray(vert32, np.float32)
    t = np.ascontiguousarray(tri, np.int32)
    o = np.ascontiguousarray(org, np.float32)
    d = np.ascontiguousarray(dirs, np.float32)
    hit = np.zeros(o.shape[0], np.uint8)
    lib().orc_occluded32(_p(v, _F), v.shape[0], _p(t, _I), t.shape[0], _p(o, _F), _p(d, _F),
                         o.shape[0], _F(tfar), int(use_bvh), _p(hit, ctypes.c_ubyte))
    return hit.astype(bool)


def occluded64(vert, tri, org, dirs, tfar):
    v = np.ascontiguousarray(vert, np.float64)
    t = np.ascontiguousarray(tri, np.int32)
    o = np.ascontiguousarray(org, np.float64)
    d = np.ascontiguousarray(dirs, np.float64)
    hit = np.zeros(o.shape[0], np.uint8)
    margin = np.zeros(o.shape[0], np.float64)
    lib().orc_occluded64(_p(v, _D), v.shape[0], _p(t, _I), t.shape[0], _p(o, _D), _p(d, _D),
                         o.shape[0], _D(tfar), _p(hit, ctypes.c_ubyte), _p(margin, _D))
    return hit.astype(bool), margin


def true_horizon(dom, cells, num_hori=24, grid_type=1, dist_search=40000.0,
                 ray_org_elev=0.2, refine_factor=1, nthreads=0, brute=False):
    """FP64 half-plane maximum elevation angle [rad], shape (len(cells), L).
    brute=True evaluates every triangle at every azimuth; the default visits
    only the triangles that can meet each azimuth's half-plane within
    dist_search (same bits)."""
    clon, clat, hsurf, vlon, vlat, cov = _inputs(dom)
    cells = np.ascontiguousarray(cells, np.int32)
    L = num_hori * refine_factor
    out = np.empty((cells.size, L), np.float64)
    lib().orc_true_horizon(_p(clon, _D), _p(clat, _D), _p(hsurf, _D), _p(vlon, _D),
                           _p(vlat, _D), _p(cov, _I), clon.size, vlon.size, num_hori,
                           grid_type, _D(dist_search), _D(ray_org_elev), refine_factor,
                           _p(cells, _I), cells.size, _p(out, _D), nthreads, int(brute))
    return out
