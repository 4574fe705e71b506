"""Device prologue (hz_options.prologue = HZ_PROLOGUE_DEVICE) against the
host prologue and the oracle.

* the device's correctly rounded sin / cos equal the exact-decimal fixture
  bit for bit (the host evaluation is checked in test_crtrig.py);
* the scene a domain traces (FP32 vertices, triangles, FP32 ray origins)
  equals the host prologue's bit for bit on C1 (both grid types) and C2;
  the FP64 unit vectors of the frames agree to 1e-15 (they are derived from
  sin / cos values that glibc rounds differently on ~0.2% of arguments);
* horizons computed from the device scene equal the oracle's (C1) and the
  host-prologue run's (C2) bit for bit; slope / aspect and the learned
  mode's descriptors agree with the host-prologue domain's;
* an argument outside the reduction's exact range is rejected.
"""

import ctypes
import os

import numpy as np
import pytest

import oracle
from horayzon_extpar_amd import _lib, synthetic
from horayzon_extpar_amd._lib import HzError
from horayzon_extpar_amd.horizon_svf import ResidentDomain, horizon_svf_comp_py, scene_export

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SVF_TOL = 1e-12


def test_device_sincos_matches_fixture():
    print("case device sincos fixture", flush=True)
    z = np.load(os.path.join(HERE, "golden", "sincos_cr.npz"), allow_pickle=False)
    x = np.ascontiguousarray(z["x"])
    s, c = np.empty_like(x), np.empty_like(x)
    p = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert _lib.lib().hz_sincos(p(x), x.size, p(s), p(c), 1) == 0
    assert np.array_equal(s.view(np.int64), z["sin"].view(np.int64))
    assert np.array_equal(c.view(np.int64), z["cos"].view(np.int64))


def _compare_scene(dom_in, grid_type, label):
    host = scene_export(*dom_in.args(), grid_type=grid_type)
    with ResidentDomain(*dom_in.args(), grid_type=grid_type, ray_org_elev=0.2,
                        prologue="device") as dom:
        dev = dom.scene()
        st = dom.create_stats
        assert (st.lon0, st.lat0) == tuple(host["lon0_lat0"])
    assert dev["tri"].shape == host["tri"].shape
    assert np.array_equal(dev["tri"], host["tri"]), label
    assert np.array_equal(dev["vert32"].view(np.int32), host["vert32"].view(np.int32)), label
    assert np.array_equal(dev["org"].view(np.int32), host["org"].view(np.int32)), label
    # unit vectors: a last-bit change of a sin / cos moves the ECEF point by
    # ~1e-9 m, i.e. the unit vectors by ~2e-16 (many ulp of a small component)
    dn = np.abs(dev["frame_n"] - host["frame_n"])
    dN = np.abs(dev["frame_north"] - host["frame_north"])
    print(f"  {label}: {dev['tri'].shape[0]} triangles, frame components differing from the "
          f"host's: n {int((dn > 0).sum())}, north {int((dN > 0).sum())} of {dn.size} "
          f"(max {dn.max():.1e} / {dN.max():.1e})", flush=True)
    assert dn.max() <= 1e-15 and dN.max() <= 1e-15


@pytest.mark.parametrize("grid_type", [0, 1])
def test_scene_c1(c1, grid_type):
    print(f"case device scene c1 gt{grid_type}", flush=True)
    _compare_scene(c1, grid_type, f"c1 gt{grid_type}")


def test_scene_c2(c2):
    print("case device scene c2", flush=True)
    _compare_scene(c2, 1, "c2 gt1")


@pytest.mark.parametrize("grid_type", [0, 1])
def test_horizon_c1_device_prologue(c1, grid_type):
    print(f"case device prologue horizons c1 gt{grid_type}", flush=True)
    st = {}
    h, s = horizon_svf_comp_py(*c1.args(), 24, grid_type, 40000.0, 0.2, 1, 2,
                               prologue="device", stats=st)
    ho, so, _ = oracle.horizon_svf(c1, grid_type=grid_type)
    assert np.array_equal(h, ho)
    assert np.max(np.abs(s - so)) <= SVF_TOL


def test_horizon_c2_device_equals_host_prologue(c2):
    print("case device prologue horizons c2", flush=True)
    with ResidentDomain(*c2.args(), grid_type=1, ray_org_elev=0.2, prologue="device") as dom:
        st_dev = dom.create_stats
        dom.run(24, 40000.0, 1, 2)
        hd, sd = dom.fetch()
    with ResidentDomain(*c2.args(), grid_type=1, ray_org_elev=0.2) as dom:
        st_host = dom.create_stats
        dom.run(24, 40000.0, 1, 2)
        hh, sh = dom.fetch()
    print(f"  c2 create: device prologue {st_dev.t_prologue_ms:.1f} + {st_dev.t_upload_ms:.1f} ms, "
          f"host prologue {st_host.t_prologue_ms:.1f} + upload {st_host.t_upload_ms:.1f} ms",
          flush=True)
    assert np.array_equal(hd, hh)
    assert np.max(np.abs(sd - sh)) <= SVF_TOL


def test_slope_aspect_device_prologue(c1):
    print("case device prologue slope/aspect", flush=True)
    out = []
    for where in ("host", "device"):
        with ResidentDomain(*c1.args(), grid_type=1, ray_org_elev=0.2, prologue=where) as dom:
            out.append(dom.slope_aspect())
    (sh, ah), (sd, ad) = out
    assert np.max(np.abs(sh - sd)) <= 1e-9
    # aspect is the azimuth of the normal's horizontal part, ill-conditioned
    # on flat cells: a 2e-16 change of the normal turns it by ~2e-16 / sin(slope)
    m = sh > 0.01
    da = np.abs((ah[m] - ad[m] + 180.0) % 360.0 - 180.0)
    assert m.sum() > 1000 and np.max(da) <= 1e-8


def test_descriptors_device_prologue(c1):
    # the learned mode's raster uses the scene's x / y extent: same scene and
    # extent give the same descriptors
    print("case device prologue descriptors", flush=True)
    out = []
    for where in ("host", "device"):
        with ResidentDomain(*c1.args(), grid_type=1, ray_org_elev=0.2, prologue=where) as dom:
            out.append(dom.descriptors())
    assert np.array_equal(out[0], out[1])


def test_out_of_range_longitude_rejected():
    print("case device prologue range check", flush=True)
    d = synthetic.lattice_domain(8.0, 8.0, 2.0)
    clon = d.clon.copy()
    clon[0] = 2.0 ** 21
    with pytest.raises(HzError) as e:
        ResidentDomain(clon, d.clat, d.hsurf, d.vlon, d.vlat, d.cells_of_vertex,
                       prologue="device")
    assert e.value.code == _lib.HZ_EINVAL
