"""ICON grid ingestion, EXTPAR output and the border-distance diagnostic
(host side; reference verification.py:82-106, :202-203, functions.py:58-97).

* grid files round-trip exactly (NetCDF classic and .npz), with the 1-based /
  -1 fill layout on disk and 0-based / -2 fill in memory;
* vertex_of_cell is the inverse of cells_of_vertex;
* NetCDF-4 (HDF5) files fail with a clear message;
* HORIZON / SKYVIEW are written as float32;
* distance_to_border equals a brute-force nearest border vertex.
"""

import numpy as np
import pytest

from horayzon_extpar_amd import diagnostics, icon_grid, synthetic


@pytest.fixture(scope="module")
def small():
    return synthetic.config1(size_km=30.0)


@pytest.mark.parametrize("ext", [".nc", ".npz"])
def test_grid_round_trip(tmp_path, small, ext):
    g = icon_grid.grid_from_domain(small)
    g.attrs = {"grid_root": 2, "grid_level": 9}
    path = str(tmp_path / ("grid" + ext))
    icon_grid.write_grid(path, g)
    r = icon_grid.read_grid(path)
    for name in ("clon", "clat", "vlon", "vlat", "vertex_of_cell", "cells_of_vertex"):
        assert np.array_equal(getattr(r, name), getattr(g, name)), name
    assert r.cells_of_vertex.min() == -2
    assert int(r.attrs["grid_level"]) == 9
    assert r.resolution_m == pytest.approx(5050e3 / (2 * 2 ** 9))


def test_on_disk_layout_is_one_based(tmp_path, small):
    from scipy.io import netcdf_file
    g = icon_grid.grid_from_domain(small)
    path = str(tmp_path / "grid.nc")
    icon_grid.write_grid(path, g)
    with netcdf_file(path, "r", mmap=False) as f:
        cov = np.array(f.variables["cells_of_vertex"].data)
        voc = np.array(f.variables["vertex_of_cell"].data)
    assert cov.min() == -1 and voc.min() == 1
    assert np.array_equal(cov - 1, g.cells_of_vertex)


def test_vertex_of_cell_inverts_cells_of_vertex(small):
    voc = icon_grid.vertex_of_cell_from_cells_of_vertex(small.cells_of_vertex, small.ncell)
    assert voc.shape == (3, small.ncell)
    cov = small.cells_of_vertex
    for c in range(0, small.ncell, 97):
        for v in voc[:, c]:
            assert c in cov[:, v]
    # each (vertex, cell) incidence appears once in each table
    assert (cov >= 0).sum() == 3 * small.ncell


def test_topography_and_size_check(tmp_path, small):
    path = str(tmp_path / "extpar.nc")
    icon_grid.write_topography(path, small.hsurf)
    h = icon_grid.read_topography(path, small.ncell)
    assert np.array_equal(h, small.hsurf.astype(np.float32).astype(np.float64))
    with pytest.raises(ValueError, match="Inconsistent number of cells"):
        icon_grid.read_topography(path, small.ncell + 1)


def test_hdf5_files_fail_clearly(tmp_path):
    path = tmp_path / "grid4.nc"
    path.write_bytes(b"\x89HDF\r\n\x1a\n" + b"\0" * 64)
    with pytest.raises(ValueError, match="NetCDF-4"):
        icon_grid.read_grid(str(path))


@pytest.mark.parametrize("ext", [".nc", ".npz"])
def test_horizon_writer_float32(tmp_path, ext):
    rng = np.random.default_rng(3)
    hz = rng.uniform(0, 30, (24, 500))
    sv = rng.uniform(0.5, 1, 500)
    path = str(tmp_path / ("out" + ext))
    icon_grid.write_horizon(path, hz, sv, num_hori=24)
    h2, s2 = icon_grid.read_horizon(path)
    assert h2.dtype == np.float32 and s2.dtype == np.float32
    assert np.array_equal(h2, hz.astype(np.float32)) and np.array_equal(s2, sv.astype(np.float32))


def test_distance_to_border_brute_force(small):
    d = diagnostics.distance_to_border(small.clon, small.clat, small.vlon, small.vlat,
                                       small.cells_of_vertex)
    border = np.any(small.cells_of_vertex == -2, axis=0)
    p = diagnostics._xyz(small.vlon[border], small.vlat[border], diagnostics.RAD_E_KM)
    q = diagnostics._xyz(small.clon, small.clat, diagnostics.RAD_E_KM)
    ref = np.sqrt(((q[:, None, :] - p[None, :, :]) ** 2).sum(-1)).min(1)
    assert np.allclose(d, ref, rtol=0, atol=1e-9)
    # the outer cells have a border vertex as a corner: circumradius 2 km / sqrt(3);
    # the centre of the 30 km box is about half its width away
    assert d.min() == pytest.approx(2.0 / np.sqrt(3.0), rel=1e-3)
    assert 10.0 < d.max() < 16.0


@pytest.mark.gpu
def test_from_files_equals_arrays(tmp_path):
    # read grid + topography files, run, write float32 fields: the same numbers
    # as the array call (topography stored as float32 on disk, as EXTPAR does)
    from horayzon_extpar_amd.horizon_svf import horizon_svf_comp_py
    print("case icon files c1", flush=True)
    d = synthetic.config1(size_km=60.0)
    g = icon_grid.grid_from_domain(d)
    gp, tp, op = (str(tmp_path / n) for n in ("grid.nc", "extpar.nc", "out.nc"))
    icon_grid.write_grid(gp, g)
    icon_grid.write_topography(tp, d.hsurf)
    h, s = icon_grid.horizon_svf_from_files(gp, tp, out_file=op)
    h0, s0 = horizon_svf_comp_py(d.clon, d.clat, d.hsurf.astype(np.float32).astype(np.float64),
                                 d.vlon, d.vlat, d.cells_of_vertex, 24, 1, 40000.0, 0.2, 1, 2)
    assert np.array_equal(h, h0) and np.array_equal(s, s0)
    h2, s2 = icon_grid.read_horizon(op)
    assert np.array_equal(h2, h0.astype(np.float32)) and np.array_equal(s2, s0.astype(np.float32))
