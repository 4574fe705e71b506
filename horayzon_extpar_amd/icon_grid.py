"""ICON grid ingestion and EXTPAR-style output: the steps on either side of
the horizon path.

Reading follows what the reference's verification script does before
calling the path (verification.py:82-104, development.py:186-201):

* the grid file holds ``clon, clat, vlon, vlat`` [rad] and the connectivity
  ``vertex_of_cell`` (3, ncell) and ``cells_of_vertex`` (6, nvert), 1-based
  with -1 fill; the path wants 0-based indices, so the fill becomes -2
  ("no adjacent cell"), as the reference's ``- 1`` makes it;
* the global attributes ``grid_root`` (n) and ``grid_level`` (k) give the
  nominal resolution 5050 km / (n 2^k) (verification.py:91-93);
* the EXTPAR file holds ``topography_c`` (cell elevation [m]).

Writing follows what the verification script hands back to EXTPAR
(verification.py:202-203): ``HORIZON`` (nhori, ncell) and ``SKYVIEW``
(ncell) as float32.

Files are NetCDF classic (CDF-1 / 64-bit offset), read and written with
scipy.io; the same variables can also live in a ``.npz`` fixture.  NetCDF-4
files are HDF5 containers: no HDF5 reader is importable in this
environment, so they raise a clear error (convert them to the classic
format, e.g. ``nccopy -k classic``).
"""

from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

from .horizon_svf import horizon_svf_comp_py

_HDF5_MAGIC = b"\x89HDF\r\n\x1a\n"


@dataclass
class IconGrid:
    """ICON grid arrays in the units and index base the path expects."""

    clon: np.ndarray             # f64 [ncell] rad
    clat: np.ndarray             # f64 [ncell] rad
    vlon: np.ndarray             # f64 [nvert] rad
    vlat: np.ndarray             # f64 [nvert] rad
    vertex_of_cell: np.ndarray   # i32 [3, ncell], 0-based
    cells_of_vertex: np.ndarray  # i32 [6, nvert], 0-based, -2 = no cell
    attrs: dict = field(default_factory=dict)

    @property
    def ncell(self) -> int:
        return int(self.clon.size)

    @property
    def nvert(self) -> int:
        return int(self.vlon.size)

    @property
    def resolution_m(self) -> float | None:
        """Nominal resolution from grid_root / grid_level (None if absent)."""
        if "grid_root" in self.attrs and "grid_level" in self.attrs:
            return grid_resolution_m(int(self.attrs["grid_root"]), int(self.attrs["grid_level"]))
        return None


def grid_resolution_m(grid_root: int, grid_level: int) -> float:
    """Nominal ICON resolution R2B-k: 5050 km / (n 2^k), in metres
    (verification.py:91-93)."""
    return 5050.0 / (grid_root * 2 ** grid_level) * 1000.0


def _check_not_hdf5(path: str):
    with open(path, "rb") as f:
        if f.read(8) == _HDF5_MAGIC:
            raise ValueError(
                f"{path} is a NetCDF-4 (HDF5) file; no HDF5 reader is available here. "
                "Convert it to NetCDF classic (nccopy -k classic in out) or to .npz.")


def _read_vars(path: str, names, attrs=()):
    """Variables (as numpy arrays) and global attributes of a NetCDF
    classic or .npz file."""
    if path.endswith(".npz"):
        z = np.load(path, allow_pickle=False)
        out = {n: np.asarray(z[n]) for n in names if n in z.files}
        at = {a: z[a].item() for a in attrs if a in z.files}
    else:
        from scipy.io import netcdf_file
        _check_not_hdf5(path)
        with netcdf_file(path, "r", mmap=False) as f:
            # NetCDF stores big-endian values: return native byte order
            out = {n: np.array(f.variables[n].data).astype(
                       f.variables[n].data.dtype.newbyteorder("="))
                   for n in names if n in f.variables}
            at = {}
            for a in attrs:
                if hasattr(f, a):
                    v = getattr(f, a)
                    at[a] = v.decode() if isinstance(v, bytes) else v
    missing = [n for n in names if n not in out]
    if missing:
        raise KeyError(f"{path}: missing variables {missing}")
    return out, at


def read_grid(path: str) -> IconGrid:
    """Read an ICON grid file (verification.py:82-90): coordinates as float64
    radians, connectivity converted from 1-based (-1 fill) to 0-based (-2
    fill)."""
    v, at = _read_vars(path, ("clon", "clat", "vlon", "vlat", "vertex_of_cell",
                              "cells_of_vertex"), ("grid_root", "grid_level"))
    voc = np.asarray(v["vertex_of_cell"], np.int32) - 1
    cov = np.asarray(v["cells_of_vertex"], np.int32) - 1
    if voc.shape[0] != 3 or cov.shape[0] != 6:
        raise ValueError(f"{path}: vertex_of_cell must be (3, ncell) and cells_of_vertex "
                         f"(6, nvert); got {voc.shape} and {cov.shape}")
    g = IconGrid(clon=np.ascontiguousarray(v["clon"], np.float64),
                 clat=np.ascontiguousarray(v["clat"], np.float64),
                 vlon=np.ascontiguousarray(v["vlon"], np.float64),
                 vlat=np.ascontiguousarray(v["vlat"], np.float64),
                 vertex_of_cell=np.ascontiguousarray(voc),
                 cells_of_vertex=np.ascontiguousarray(cov), attrs=at)
    if g.vertex_of_cell.shape[1] != g.ncell or g.cells_of_vertex.shape[1] != g.nvert:
        raise ValueError(f"{path}: connectivity does not match the coordinate sizes")
    return g


def read_topography(path: str, ncell: int | None = None) -> np.ndarray:
    """Cell elevation ``topography_c`` [m] of an EXTPAR file, squeezed to 1-D
    float64 (verification.py:97-99); checked against the grid size as the
    verification script does (:105-106)."""
    v, _ = _read_vars(path, ("topography_c",))
    h = np.ascontiguousarray(np.squeeze(v["topography_c"]), np.float64)
    if ncell is not None and h.size != ncell:
        raise ValueError("Inconsistent number of cells in grid and topography")
    return h


def vertex_of_cell_from_cells_of_vertex(cells_of_vertex: np.ndarray, ncell: int) -> np.ndarray:
    """Invert the vertex -> cells adjacency (0-based, -2 fill) into the three
    vertices of every cell, in increasing vertex order."""
    cov = np.asarray(cells_of_vertex)
    v = np.broadcast_to(np.arange(cov.shape[1]), cov.shape)
    m = cov >= 0
    cells, verts = cov[m], v[m]
    order = np.lexsort((verts, cells))
    cells, verts = cells[order], verts[order]
    counts = np.bincount(cells, minlength=ncell)
    if np.any(counts != 3):
        raise ValueError("every cell must be a corner-sharing triangle of exactly 3 vertices")
    return np.ascontiguousarray(verts.reshape(ncell, 3).T.astype(np.int32))


def write_grid(path: str, grid: IconGrid):
    """Write the grid in the ICON file layout (1-based, -1 fill) as NetCDF
    classic, or as .npz when path ends with .npz (test fixtures)."""
    voc = (grid.vertex_of_cell + 1).astype(np.int32)
    cov = (grid.cells_of_vertex + 1).astype(np.int32)
    if path.endswith(".npz"):
        np.savez(path, clon=grid.clon, clat=grid.clat, vlon=grid.vlon, vlat=grid.vlat,
                 vertex_of_cell=voc, cells_of_vertex=cov,
                 **{k: np.asarray(v) for k, v in grid.attrs.items()})
        return
    from scipy.io import netcdf_file
    with netcdf_file(path, "w", version=2) as f:
        f.createDimension("cell", grid.ncell)
        f.createDimension("vertex", grid.nvert)
        f.createDimension("nv", 3)
        f.createDimension("ne", 6)
        for name, dim, data in (("clon", "cell", grid.clon), ("clat", "cell", grid.clat),
                                ("vlon", "vertex", grid.vlon), ("vlat", "vertex", grid.vlat)):
            var = f.createVariable(name, "d", (dim,))
            var[:] = data
            var.units = "radian"
        var = f.createVariable("vertex_of_cell", "i", ("nv", "cell"))
        var[:] = voc
        var = f.createVariable("cells_of_vertex", "i", ("ne", "vertex"))
        var[:] = cov
        for k, v in grid.attrs.items():
            setattr(f, k, v)


def write_topography(path: str, hsurf: np.ndarray):
    """EXTPAR-like file with ``topography_c`` (1, ncell) [m] (test fixtures)."""
    if path.endswith(".npz"):
        np.savez(path, topography_c=np.asarray(hsurf, np.float64)[None, :])
        return
    from scipy.io import netcdf_file
    with netcdf_file(path, "w", version=2) as f:
        f.createDimension("time", 1)
        f.createDimension("cell", hsurf.size)
        var = f.createVariable("topography_c", "f", ("time", "cell"))
        var[:] = np.asarray(hsurf, np.float32)[None, :]
        var.units = "m"


def write_horizon(path: str, horizon: np.ndarray, skyview: np.ndarray, **attrs):
    """EXTPAR output fields: ``HORIZON`` (nhori, ncell) [deg] and ``SKYVIEW``
    (ncell) [-] as float32 (verification.py:202-203), NetCDF classic or .npz."""
    hz = np.asarray(horizon, np.float32)
    sv = np.asarray(skyview, np.float32)
    if hz.ndim != 2 or hz.shape[1] != sv.size:
        raise ValueError("horizon must be (nhori, ncell) and skyview (ncell,)")
    if path.endswith(".npz"):
        np.savez(path, HORIZON=hz, SKYVIEW=sv)
        return
    from scipy.io import netcdf_file
    with netcdf_file(path, "w", version=2) as f:
        f.createDimension("nhori", hz.shape[0])
        f.createDimension("cell", hz.shape[1])
        var = f.createVariable("HORIZON", "f", ("nhori", "cell"))
        var[:] = hz
        var.units = "deg"
        var.long_name = "terrain horizon angle"
        var = f.createVariable("SKYVIEW", "f", ("cell",))
        var[:] = sv
        var.units = "1"
        var.long_name = "sky view factor"
        for k, v in attrs.items():
            setattr(f, k, v)


def read_horizon(path: str):
    """(HORIZON, SKYVIEW) float32 arrays of an output file."""
    v, _ = _read_vars(path, ("HORIZON", "SKYVIEW"))
    return np.squeeze(v["HORIZON"]), np.squeeze(v["SKYVIEW"])


def horizon_svf_from_files(grid_file: str, topo_file: str, num_hori: int = 24,
                           grid_type: int = 1, dist_search: float = 40000.0,
                           ray_org_elev: float = 0.2, refine_factor: int = 1,
                           svf_type: int = 2, out_file: str | None = None, **opts):
    """Read an ICON grid and its EXTPAR topography, run the path and (if
    out_file is given) write the float32 EXTPAR fields; returns the float64
    (horizon, skyview).  Defaults are the verification
    script's (verification.py:151-155)."""
    g = read_grid(grid_file)
    hsurf = read_topography(topo_file, g.ncell)
    horizon, skyview = horizon_svf_comp_py(g.clon, g.clat, hsurf, g.vlon, g.vlat,
                                           g.cells_of_vertex, num_hori, grid_type,
                                           dist_search, ray_org_elev, refine_factor, svf_type,
                                           **opts)
    if out_file:
        write_horizon(out_file, horizon, skyview, num_hori=num_hori, grid_type=grid_type,
                      dist_search=dist_search, refine_factor=refine_factor,
                      svf_type=svf_type, source=os.path.basename(grid_file))
    return horizon, skyview


def grid_from_domain(dom) -> IconGrid:
    """IconGrid of a synthetic.Domain (its connectivity completed with
    vertex_of_cell)."""
    return IconGrid(clon=dom.clon, clat=dom.clat, vlon=dom.vlon, vlat=dom.vlat,
                    vertex_of_cell=vertex_of_cell_from_cells_of_vertex(dom.cells_of_vertex,
                                                                       dom.ncell),
                    cells_of_vertex=np.ascontiguousarray(dom.cells_of_vertex, np.int32),
                    attrs={})
