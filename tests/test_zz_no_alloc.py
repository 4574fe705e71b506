"""GPU: hz_walk_noalloc, the reference walk with no device allocation.

It runs the same kernels as hz_domain_run with search_mode 0 over a scene
the caller holds in device memory (the hz_scene_export layout), with every
intermediate buffer in the library's static scratch array.  Its horizons and
sky view factors equal hz_domain_run's bit for bit, on two calls in a row
where the second domain is smaller than the first (the array keeps the
first call's contents), and it refuses scenes beyond its stated limits.
"""

import ctypes

import numpy as np
import pytest

from horayzon_extpar_amd import _lib, synthetic
from horayzon_extpar_amd.horizon_svf import ResidentDomain, scene_export

pytestmark = pytest.mark.gpu

MAX_CELLS = 65536
MAX_TRIANGLES = 196608


def _domain_run(dom, nh, rf, svf):
    with ResidentDomain(*dom.args(), grid_type=1, ray_org_elev=0.2) as rd:
        rd.run(nh, 40000.0, rf, svf, search_mode=0)
        return rd.fetch()


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime libhorayzon.so is
    linked against (the caller's own allocations; a second HIP runtime in
    the process, such as the one torch bundles, cannot share the device)."""

    def __init__(self):
        _lib.lib()
        self.rt = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                      ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.held = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), max(a.nbytes, 1)) == 0
        self.held.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p

    def download(self, p, a):
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(a.ctypes.data, p, a.nbytes, 2) == 0
        return a

    def free(self):
        for p in self.held:
            self.rt.hipFree(p)
        self.held = []


def _noalloc(dom, nh, rf, svf):
    sc = scene_export(*dom.args(), grid_type=1, ray_org_elev=0.2)
    hip = _Hip()
    try:
        dev = {k: hip.upload(sc[k]) for k in ("vert32", "tri", "frame_n", "frame_north", "org")}
        n = dom.ncell
        nan_h = np.full((nh, n), np.nan)
        nan_s = np.full(n, np.nan)
        dh, ds = hip.upload(nan_h), hip.upload(nan_s)
        rc = _lib.lib().hz_walk_noalloc(
            dev["vert32"], sc["vert32"].shape[0], dev["tri"], sc["tri"].shape[0],
            dev["frame_n"], dev["frame_north"], dev["org"], n, nh, 40000.0, rf, svf, dh, ds,
            None)
        _lib.check(rc)
        h = hip.download(dh, np.empty((nh, n)))
        s = hip.download(ds, np.empty(n))
    finally:
        hip.free()
    return h, s, sc["tri"].shape[0]


def test_noalloc_equals_domain_run_twice():
    cases = [(synthetic.config1(), 24, 1, 2), (synthetic.config1(seed=7, size_km=60.0), 12, 3, 1)]
    for dom, nh, rf, svf in cases:
        print(f"case noalloc ncell={dom.ncell} nhori={nh} rf={rf} svf={svf}", flush=True)
        h0, s0 = _domain_run(dom, nh, rf, svf)
        h1, s1, ntri = _noalloc(dom, nh, rf, svf)
        assert ntri <= MAX_TRIANGLES and dom.ncell <= MAX_CELLS
        assert np.array_equal(h0, h1), np.nanmax(np.abs(h0 - h1))
        assert np.array_equal(s0, s1)


def test_noalloc_refuses_beyond_limits():
    print("case noalloc limits", flush=True)
    L = _lib.lib()
    p = ctypes.c_void_p(16)  # never dereferenced: the size check comes first
    rc = L.hz_walk_noalloc(p, 3, p, 1, p, p, p, MAX_CELLS + 1, 24, 40000.0, 1, 2, p, p, None)
    assert rc == _lib.HZ_ENOMEM and "at most" in _lib.last_error()
    rc = L.hz_walk_noalloc(p, 3, p, MAX_TRIANGLES + 1, p, p, p, 100, 24, 40000.0, 1, 2, p, p,
                           None)
    assert rc == _lib.HZ_ENOMEM
    rc = L.hz_walk_noalloc(p, 3, p, 1, p, p, p, 100, 3, 40000.0, 1, 2, p, p, None)
    assert rc == _lib.HZ_EINVAL
