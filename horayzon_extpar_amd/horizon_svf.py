"""Drop-in Python surface of the reference binding (horizon_svf.pyx).

``horizon_svf_comp_py`` takes the same positional arguments, applies the
same checks in the same order with the same exception types and messages
(horizon_svf.pyx:79-103), allocates the outputs the same way (:105-109) and
returns ``(horizon[num_hori, ncell], skyview[ncell])`` in float64.  The work
is done by libhorayzon.so on the GPU.

Differences, all of them supersets:
* 1-D inputs are made contiguous (the reference passes ``&x[0]`` of a
  possibly strided buffer, :115-117, which reads wrong elements);
* a failing device call raises ``HzError`` (RuntimeError) instead of
  printing and returning garbage;
* keyword-only ``search_mode``, ``leaf_size``, ``block_size``,
  ``collect_counters``, ``trace_kernel``, ``refill``, ``bvh_morton``,
  ``bvh_width``, ``prologue``, ``whole_cells``, ``certify``, ``cache`` and
  ``stats`` expose the library options (``hz_options`` in include/horayzon.h).  The
  default ``search_mode=0`` is the reference walk; the lattice searches
  (1 climb, 2 learned start) equal it with the default ``certify=1``.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import HzStats, check, default_options

_F64 = np.float64
_I32 = np.int32


def _check_buffer(name, a, dtype, ndim, cname):
    # mirrors Cython's typed-buffer argument checks of the reference signature
    # (horizon_svf.pyx:15-20)
    if not isinstance(a, np.ndarray):
        raise TypeError(f"Argument '{name}' has incorrect type (expected numpy.ndarray, "
                        f"got {type(a).__name__})")
    if a.dtype != dtype:
        raise ValueError(f"Buffer dtype mismatch, expected '{cname}' but got "
                         f"'{a.dtype.name}'")
    if a.ndim != ndim:
        raise ValueError(f"Buffer has wrong number of dimensions (expected {ndim}, "
                         f"got {a.ndim})")


def _check_int(name, v):
    if isinstance(v, (bool, np.bool_)) or not isinstance(v, (int, np.integer)):
        raise TypeError(f"'{type(v).__name__}' object cannot be interpreted as an integer "
                        f"(argument '{name}')")
    return int(v)


def _ptr(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def validate(clon, clat, hsurf, vlon, vlat, cells_of_vertex, num_hori, grid_type,
             dist_search, ray_org_elev, refine_factor, svf_type):
    """Argument checks of horizon_svf.pyx:15-26 and :79-103, in order."""
    for name, a in (("clon", clon), ("clat", clat), ("hsurf", hsurf), ("vlon", vlon),
                    ("vlat", vlat)):
        _check_buffer(name, a, _F64, 1, "float64_t")
    _check_buffer("cells_of_vertex", cells_of_vertex, _I32, 2, "int32_t")
    num_hori = _check_int("num_hori", num_hori)
    grid_type = _check_int("grid_type", grid_type)
    dist_search = float(dist_search)
    ray_org_elev = float(ray_org_elev)
    refine_factor = _check_int("refine_factor", refine_factor)
    svf_type = _check_int("svf_type", svf_type)
    if (clon.size != clat.size) or (clat.size != hsurf.size):
        raise ValueError("Inconsistent lengths of input arrays 'clon', 'clat' and 'hsurf'")
    if vlon.size != vlat.size:
        raise ValueError("Inconsistent lengths of input arrays 'vlon' and 'vlat'")
    if cells_of_vertex.shape[0] != 6:
        raise ValueError("First dimension of 'cells_of_vertex' must have length 6")
    # every entry in [0, ncell - 1] or -2 (the same test as the reference's
    # elementwise one, as three reductions)
    if cells_of_vertex.size and not (cells_of_vertex.min() >= -2
                                     and cells_of_vertex.max() <= clon.size - 1
                                     and not (cells_of_vertex == -1).any()):
        raise ValueError("Indices of 'cells_of_vertex' out of range")
    if (num_hori < 4) or (num_hori > 1440):
        raise ValueError("'num_hori' must be in the range [4, 1440]")
    if (grid_type < 0) or (grid_type > 1):
        raise ValueError("'grid_type' must be in the range [0, 1]")
    if (dist_search < 1_000.0) or (dist_search > 500_000.0):
        raise ValueError("'dist_search' must be in the range [1'000, 500'000] km")
    if ray_org_elev < 0.1:
        raise TypeError("Minimal allowed value for 'ray_org_elev' is 0.1 m")
    if (refine_factor < 1) or (refine_factor > 50):
        raise ValueError("'refine_factor' must be in the range [1, 50]")
    if (svf_type < 0) or (svf_type > 2):
        raise ValueError("'svf_type' must be in the range [0, 2]")
    # the reference takes &x[0] of every buffer (:115-117): empty buffers raise
    for name, a in (("clon", clon), ("vlon", vlon)):
        if a.size == 0:
            raise IndexError(f"Out of bounds on buffer access (axis 0) ('{name}')")
    if cells_of_vertex.shape[1] == 0:
        raise IndexError("Out of bounds on buffer access (axis 1) ('cells_of_vertex')")
    return num_hori, grid_type, dist_search, ray_org_elev, refine_factor, svf_type


def horizon_svf_comp_py(clon, clat, hsurf, vlon, vlat, cells_of_vertex, num_hori,
                        grid_type, dist_search, ray_org_elev, refine_factor, svf_type, *,
                        search_mode=None, leaf_size=None, block_size=None,
                        collect_counters=None, trace_kernel=None, refill=None,
                        bvh_morton=None, bvh_width=None, prologue=None, whole_cells=None,
                        certify=None, cache=None, return_slope_aspect=False, stats=None):
    """Compute the terrain horizon [deg] and sky view factor [-].

    Same contract as the reference binding (horizon_svf.pyx:15-124):
    clon, clat, hsurf, vlon, vlat are float64 1-D; cells_of_vertex is int32
    (6, num_vertex) with 0-based cell indices and -2 fill.  Returns
    ``horizon`` (num_hori, ncell) and ``skyview`` (ncell,), float64.
    ``stats``: optional dict that receives the library statistics.
    ``return_slope_aspect=True`` appends the terrain slope and aspect [deg]
    per cell (an extension; the reference returns only the two arrays).
    ``prologue="device"`` builds the mesh, ENU scene and frames on the GPU
    (default "host": the reference's values bit for bit).
    """
    (num_hori, grid_type, dist_search, ray_org_elev, refine_factor,
     svf_type) = validate(clon, clat, hsurf, vlon, vlat, cells_of_vertex, num_hori,
                          grid_type, dist_search, ray_org_elev, refine_factor, svf_type)
    horizon = np.empty((num_hori, clon.size), dtype=np.float64)
    skyview = np.empty(clon.size, dtype=np.float64)
    cells_of_vertex = np.ascontiguousarray(cells_of_vertex)
    clon, clat, hsurf, vlon, vlat = (np.ascontiguousarray(x) for x in
                                     (clon, clat, hsurf, vlon, vlat))
    opt = default_options(search_mode=search_mode, leaf_size=leaf_size,
                          block_size=block_size, collect_counters=collect_counters,
                          trace_kernel=trace_kernel, refill=refill, bvh_morton=bvh_morton,
                          bvh_width=bvh_width, prologue=prologue, whole_cells=whole_cells,
                          certify=certify, cache=cache)
    if return_slope_aspect:
        with ResidentDomain(clon, clat, hsurf, vlon, vlat, cells_of_vertex, grid_type,
                            ray_org_elev, prologue=opt.prologue) as dom:
            kw = {f: getattr(opt, f) for f in ("search_mode", "leaf_size", "block_size",
                                               "collect_counters", "trace_kernel", "refill",
                                               "bvh_morton", "bvh_width", "whole_cells",
                                               "certify")}
            st_run = dom.run(num_hori, dist_search, refine_factor, svf_type, **kw)
            horizon, skyview = dom.fetch()
            slope, aspect = dom.slope_aspect()
        if stats is not None:
            stats.update(st_run)
        return horizon, skyview, slope, aspect
    st = HzStats()
    rc = _lib.lib().horizon_svf_comp_ex(
        _ptr(clon, ctypes.c_double), _ptr(clat, ctypes.c_double),
        _ptr(hsurf, ctypes.c_double), _ptr(vlon, ctypes.c_double),
        _ptr(vlat, ctypes.c_double), _ptr(cells_of_vertex, ctypes.c_int),
        _ptr(horizon, ctypes.c_double), _ptr(skyview, ctypes.c_double),
        clon.size, vlon.size, num_hori, grid_type, dist_search, ray_org_elev,
        refine_factor, svf_type, ctypes.byref(opt), ctypes.byref(st))
    if stats is not None:
        stats.update(st.as_dict())
    check(rc)
    return horizon, skyview


class ResidentDomain:
    """A domain whose FP32 scene and cell frames stay in device memory.

    ``run`` performs one full device pass (LBVH build, cell ordering, search,
    sector mean and SVF); ``fetch`` copies the results out.  Together they
    compute what ``horizon_svf_comp_py`` computes.
    """

    def __init__(self, clon, clat, hsurf, vlon, vlat, cells_of_vertex, grid_type=1,
                 ray_org_elev=0.2, prologue="host"):
        validate(clon, clat, hsurf, vlon, vlat, cells_of_vertex, 24, grid_type, 40000.0,
                 ray_org_elev, 1, 2)
        self._arrays = tuple(np.ascontiguousarray(x) for x in
                             (clon, clat, hsurf, vlon, vlat, cells_of_vertex))
        c = self._arrays
        self.ncell = c[0].size
        self.create_stats = HzStats()
        h = ctypes.c_void_p()
        opt = default_options(prologue=prologue)
        rc = _lib.lib().hz_domain_create(
            _ptr(c[0], ctypes.c_double), _ptr(c[1], ctypes.c_double),
            _ptr(c[2], ctypes.c_double), _ptr(c[3], ctypes.c_double),
            _ptr(c[4], ctypes.c_double), _ptr(c[5], ctypes.c_int), self.ncell,
            c[3].size, int(grid_type), float(ray_org_elev), ctypes.byref(opt), ctypes.byref(h),
            ctypes.byref(self.create_stats))
        check(rc)
        self._h = h
        self.num_hori = None

    def run(self, num_hori=24, dist_search=40000.0, refine_factor=1, svf_type=2, **opts):
        st = HzStats()
        opt = default_options(**opts)
        rc = _lib.lib().hz_domain_run(self._h, int(num_hori), float(dist_search),
                                      int(refine_factor), int(svf_type), ctypes.byref(opt),
                                      ctypes.byref(st))
        check(rc)
        self.num_hori = int(num_hori)
        return st.as_dict()

    def set_predictor(self, predictor):
        """Use ``predictor`` (learned.Predictor, kept alive by the caller) in
        search_mode=2 runs."""
        self._predictor = predictor
        check(_lib.lib().hz_domain_set_predictor(self._h, predictor.handle if predictor
                                                 else None))

    def scene(self):
        """The FP32 scene and frames this domain traces, read back from the
        device: dict vert32 (nmv, 3), tri (ntri, 3), frame_n / frame_north
        (ncell, 3) FP64, org (ncell, 3) FP32 (the layout of scene_export)."""
        nt = int(self.create_stats.num_triangles)
        nv = int(self.create_stats.num_mesh_vertices)
        out = dict(vert32=np.empty((nv, 3), np.float32), tri=np.empty((nt, 3), np.int32),
                   frame_n=np.empty((self.ncell, 3)), frame_north=np.empty((self.ncell, 3)),
                   org=np.empty((self.ncell, 3), np.float32))
        check(_lib.lib().hz_domain_scene(
            self._h, _ptr(out["vert32"], ctypes.c_float), _ptr(out["tri"], ctypes.c_int),
            _ptr(out["frame_n"], ctypes.c_double), _ptr(out["frame_north"], ctypes.c_double),
            _ptr(out["org"], ctypes.c_float)))
        return out

    def descriptors(self):
        """Per-cell descriptors, BF16 bit patterns [ncell, 192]."""
        out = np.empty((self.ncell, 192), np.uint16)
        check(_lib.lib().hz_domain_descriptors(self._h, out.ctypes.data_as(
            ctypes.POINTER(ctypes.c_uint16))))
        return out

    def bvh(self, **opts):
        """Builds this domain's traversal tree with the tree options in opts
        (leaf_size, bvh_width, bvh_morton) and returns (nodes, tris): the
        128-byte node records as a uint8 array (num_nodes, 128) and the
        leaf-order triangles as float32 (num_tris, 12)."""
        opt = default_options(**opts)
        nn, nt = ctypes.c_longlong(), ctypes.c_longlong()
        L = _lib.lib()
        check(L.hz_domain_bvh(self._h, ctypes.byref(opt), None, ctypes.byref(nn), None,
                              ctypes.byref(nt)))
        nodes = np.empty((nn.value, 128), np.uint8)
        tris = np.empty((nt.value, 12), np.float32)
        check(L.hz_domain_bvh(self._h, ctypes.byref(opt), nodes.ctypes.data_as(ctypes.c_void_p),
                              ctypes.byref(nn), _ptr(tris, ctypes.c_float), ctypes.byref(nt)))
        return nodes, tris

    def slope_aspect(self):
        """Terrain slope and aspect [deg] per cell (aspect: downhill azimuth,
        clockwise from north)."""
        slope = np.empty(self.ncell, np.float64)
        aspect = np.empty(self.ncell, np.float64)
        check(_lib.lib().hz_domain_slope_aspect(self._h, _ptr(slope, ctypes.c_double),
                                                _ptr(aspect, ctypes.c_double)))
        return slope, aspect

    def predictions(self):
        """Predicted horizon [ncell, 24] (deg) of the last search_mode=2 run."""
        out = np.empty((self.ncell, 24), np.float32)
        check(_lib.lib().hz_domain_predictions(self._h, out.ctypes.data_as(
            ctypes.POINTER(ctypes.c_float))))
        return out

    def fetch(self):
        horizon = np.empty((self.num_hori, self.ncell), np.float64)
        skyview = np.empty(self.ncell, np.float64)
        st = HzStats()
        check(_lib.lib().hz_domain_fetch(self._h, _ptr(horizon, ctypes.c_double),
                                         _ptr(skyview, ctypes.c_double), ctypes.byref(st)))
        return horizon, skyview

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().hz_domain_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def scene_export(clon, clat, hsurf, vlon, vlat, cells_of_vertex, grid_type=1,
                 ray_org_elev=0.2):
    """Host-only diagnostic: the FP32 scene and cell frames the library
    traces (mesh builder + FP64 prologue).  Returns a dict of arrays."""
    a = [np.ascontiguousarray(x) for x in (clon, clat, hsurf, vlon, vlat)]
    cov = np.ascontiguousarray(cells_of_vertex, dtype=np.int32)
    L = _lib.lib()
    nt, nv = ctypes.c_longlong(), ctypes.c_longlong()
    args = [_ptr(x, ctypes.c_double) for x in a] + [_ptr(cov, ctypes.c_int), a[0].size,
                                                     a[3].size, int(grid_type),
                                                     float(ray_org_elev)]
    check(L.hz_scene_export(*args, ctypes.byref(nt), ctypes.byref(nv), None, None, None,
                            None, None, None))
    ncell = a[0].size
    out = dict(vert32=np.empty((nv.value, 3), np.float32),
               tri=np.empty((nt.value, 3), np.int32),
               frame_n=np.empty((ncell, 3), np.float64),
               frame_north=np.empty((ncell, 3), np.float64),
               org=np.empty((ncell, 3), np.float32), lon0_lat0=np.empty(2, np.float64))
    check(L.hz_scene_export(*args, ctypes.byref(nt), ctypes.byref(nv),
                            _ptr(out["vert32"], ctypes.c_float), _ptr(out["tri"], ctypes.c_int),
                            _ptr(out["frame_n"], ctypes.c_double),
                            _ptr(out["frame_north"], ctypes.c_double),
                            _ptr(out["org"], ctypes.c_float),
                            _ptr(out["lon0_lat0"], ctypes.c_double)))
    return out
