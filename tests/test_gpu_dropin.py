"""GPU: the exact drop-in symbol, as the reference's callers bind it.

* ``void horizon_svf_comp`` (mo_lradtopo_horayzon.h:4-11) called through a
  bare ctypes stub, as INTEGRATION.md's stub does: bit-identical to
  horizon_svf_comp_py on C1 (grid_type 0 and 1), and NaN in both outputs
  when the arguments are out of range;
* the same symbol through a Cython extension whose extern block has the
  signature of the reference's binding (horizon_svf.pyx:4-12) against
  include/horayzon.h (built by build()): bit-identical as well, and NaN on
  the same invalid call.
"""

import ctypes

import numpy as np
import pytest

from horayzon_extpar_amd import _lib
from horayzon_extpar_amd.horizon_svf import horizon_svf_comp_py

pytestmark = pytest.mark.gpu

DP = ctypes.POINTER(ctypes.c_double)
IP = ctypes.POINTER(ctypes.c_int)


def _stub():
    L = ctypes.CDLL(_lib.library_path())
    f = L.horizon_svf_comp
    f.argtypes = [DP] * 5 + [IP, DP, DP] + [ctypes.c_int] * 4 + [ctypes.c_double] * 2 + \
        [ctypes.c_int] * 2
    f.restype = None
    return f


def _call(f, dom, num_hori, grid_type, dist, elev, rf, svf):
    p = lambda a, t: a.ctypes.data_as(t)  # noqa: E731
    cov = np.ascontiguousarray(dom.cells_of_vertex, np.int32)
    h = np.zeros((num_hori, dom.ncell))
    s = np.zeros(dom.ncell)
    f(p(dom.clon, DP), p(dom.clat, DP), p(dom.hsurf, DP), p(dom.vlon, DP), p(dom.vlat, DP),
      p(cov, IP), p(h, DP), p(s, DP), dom.ncell, dom.vlon.size, num_hori, grid_type, dist,
      elev, rf, svf)
    return h, s


@pytest.mark.parametrize("grid_type,num_hori,rf,svf", [(1, 24, 1, 2), (0, 12, 3, 1)])
def test_void_symbol_equals_python(c1, grid_type, num_hori, rf, svf):
    print(f"case void horizon_svf_comp gt={grid_type} nhori={num_hori}", flush=True)
    h, s = _call(_stub(), c1, num_hori, grid_type, 40000.0, 0.2, rf, svf)
    h0, s0 = horizon_svf_comp_py(*c1.args(), num_hori, grid_type, 40000.0, 0.2, rf, svf)
    assert np.array_equal(h, h0) and np.array_equal(s, s0)


@pytest.mark.parametrize("bad", [dict(num_hori=3), dict(grid_type=2), dict(dist=500.0),
                                 dict(elev=0.05), dict(rf=0), dict(svf=3)])
def test_void_symbol_nan_fill(c1, bad):
    print(f"case void horizon_svf_comp invalid {bad}", flush=True)
    a = dict(num_hori=24, grid_type=1, dist=40000.0, elev=0.2, rf=1, svf=2)
    a.update(bad)
    h, s = _call(_stub(), c1, max(a["num_hori"], 1), a["grid_type"], a["dist"], a["elev"],
                 a["rf"], a["svf"])
    assert np.isnan(h).all() and np.isnan(s).all()


def test_cython_relink_equals_python(c1):
    print("case cython relink", flush=True)
    from horayzon_extpar_amd import horizon_svf_relink as m
    cov = np.ascontiguousarray(c1.cells_of_vertex, np.int32)
    h, s = m.horizon_svf_comp_relinked(c1.clon, c1.clat, c1.hsurf, c1.vlon, c1.vlat, cov,
                                       24, 1, 40000.0, 0.2, 1, 2)
    h0, s0 = horizon_svf_comp_py(*c1.args(), 24, 1, 40000.0, 0.2, 1, 2)
    assert np.array_equal(h, h0) and np.array_equal(s, s0)
    h, s = m.horizon_svf_comp_relinked(c1.clon, c1.clat, c1.hsurf, c1.vlon, c1.vlat, cov,
                                       24, 1, 40000.0, 0.2, 0, 2)
    assert np.isnan(h).all() and np.isnan(s).all()
