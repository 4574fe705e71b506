"""Host side of libhorayzon.so, CPU only: the library loads and exports every
entry point of include/horayzon.h, the host FP64 prologue equals the
oracle's bit for bit, and argument checking follows the reference binding
(horizon_svf.pyx:79-103) in order, exception type and message.  No kernel is
launched here."""

import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from horayzon_extpar_amd import _lib, synthetic
from horayzon_extpar_amd.horizon_svf import (horizon_svf_comp_py, scene_export,
                                             validate)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    txt = open(os.path.join(ROOT, "include", "horayzon.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\(", txt, flags=re.M)
    return sorted(set(n for n in names if n not in ("defined",)))


def test_exports_every_declared_symbol():
    L = _lib.lib()
    names = _declared_functions()
    assert "horizon_svf_comp" in names and "hz_domain_run" in names
    for n in names:
        assert hasattr(L, n), n
    # and the ctypes table binds exactly these
    assert sorted(s[0] for s in _lib.SIGNATURES) == names


def test_default_options():
    o = _lib.default_options()
    assert o.search_mode == 0 and o.leaf_size >= 1 and o.cell_order == 1
    assert o.whole_cells == 0 and o.prologue == 0
    # hz_options is 16 ints (include/horayzon.h)
    assert ctypes.sizeof(_lib.HzOptions) == 16 * 4
    with pytest.raises(TypeError):
        _lib.default_options(no_such_option=1)


@pytest.mark.parametrize("grid_type", [0, 1])
def test_prologue_equals_oracle(c1, grid_type):
    S = scene_export(*c1.args(), grid_type=grid_type)
    E = oracle.enu(c1, grid_type)
    _, tri = oracle.mesh(c1, grid_type)
    assert np.array_equal(S["tri"], tri)
    assert np.array_equal(S["vert32"].view(np.uint32), E["vert32"].view(np.uint32))
    assert np.array_equal(S["frame_n"], E["frame_n"])
    assert np.array_equal(S["frame_north"], E["frame_north"])
    assert np.array_equal(S["org"].view(np.uint32), E["org"].view(np.uint32))
    assert S["lon0_lat0"][0] == E["lon0"] and S["lon0_lat0"][1] == E["lat0"]


def test_prologue_equals_oracle_large(c2):
    S = scene_export(*c2.args(), grid_type=1)
    E = oracle.enu(c2, 1)
    assert np.array_equal(S["vert32"].view(np.uint32), E["vert32"].view(np.uint32))
    assert np.array_equal(S["org"].view(np.uint32), E["org"].view(np.uint32))
    assert np.array_equal(S["frame_n"], E["frame_n"])


def _args(d, **kw):
    a = dict(num_hori=24, grid_type=1, dist_search=40000.0, ray_org_elev=0.2,
             refine_factor=1, svf_type=2)
    a.update(kw)
    return (*d.args(), a["num_hori"], a["grid_type"], a["dist_search"], a["ray_org_elev"],
            a["refine_factor"], a["svf_type"])


@pytest.mark.parametrize("kw,exc,msg", [
    (dict(num_hori=3), ValueError, "'num_hori' must be in the range [4, 1440]"),
    (dict(num_hori=1441), ValueError, "'num_hori'"),
    (dict(grid_type=2), ValueError, "'grid_type' must be in the range [0, 1]"),
    (dict(dist_search=999.0), ValueError, "'dist_search' must be in the range"),
    (dict(dist_search=500001.0), ValueError, "'dist_search'"),
    (dict(ray_org_elev=0.05), TypeError, "Minimal allowed value for 'ray_org_elev' is 0.1 m"),
    (dict(refine_factor=0), ValueError, "'refine_factor' must be in the range [1, 50]"),
    (dict(refine_factor=51), ValueError, "'refine_factor'"),
    (dict(svf_type=3), ValueError, "'svf_type' must be in the range [0, 2]"),
])
def test_validation_messages(kw, exc, msg):
    d = synthetic.lattice_domain(10.0, 10.0, 2.0)
    with pytest.raises(exc) as ei:
        horizon_svf_comp_py(*_args(d, **kw))
    assert msg in str(ei.value)


def test_reference_module_name():
    # the reference's import line, verification.py: from horizon_svf import ...
    import horizon_svf
    from horayzon_extpar_amd.horizon_svf import horizon_svf_comp_py as impl
    assert horizon_svf.horizon_svf_comp_py is impl
    d = synthetic.lattice_domain(10.0, 10.0, 2.0)
    with pytest.raises(TypeError) as ei:
        horizon_svf.horizon_svf_comp_py(*_args(d, ray_org_elev=0.05))
    assert "Minimal allowed value for 'ray_org_elev' is 0.1 m" in str(ei.value)


def test_validation_order_and_arrays():
    d = synthetic.lattice_domain(10.0, 10.0, 2.0)
    a = list(_args(d))
    b = list(a)
    b[1] = b[1][:-1]
    with pytest.raises(ValueError, match="'clon', 'clat' and 'hsurf'"):
        validate(*b)
    b = list(a)
    b[4] = b[4][:-1]
    with pytest.raises(ValueError, match="'vlon' and 'vlat'"):
        validate(*b)
    b = list(a)
    b[5] = b[5][:5]
    with pytest.raises(ValueError, match="First dimension"):
        validate(*b)
    b = list(a)
    cov = b[5].copy()
    cov[0, 0] = d.ncell
    b[5] = cov
    with pytest.raises(ValueError, match="out of range"):
        validate(*b)
    cov[0, 0] = -1
    with pytest.raises(ValueError, match="out of range"):
        validate(*b)
    # inconsistent lengths win over a bad num_hori (order of the checks)
    b = list(a)
    b[2] = b[2][:-1]
    b[6] = 1
    with pytest.raises(ValueError, match="'clon', 'clat' and 'hsurf'"):
        validate(*b)
    # typed buffers
    b = list(a)
    b[0] = b[0].astype(np.float32)
    with pytest.raises(ValueError, match="Buffer dtype mismatch"):
        validate(*b)
    b = list(a)
    b[5] = b[5].astype(np.int64)
    with pytest.raises(ValueError, match="Buffer dtype mismatch"):
        validate(*b)
    b = list(a)
    b[0] = list(b[0])
    with pytest.raises(TypeError):
        validate(*b)
    b = list(a)
    b[6] = 24.0
    with pytest.raises(TypeError):
        validate(*b)


def _has_device():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


def _raw_call(d, num_hori=24, rf=1):
    L = _lib.lib()
    a = [np.ascontiguousarray(x) for x in d.args()]
    h = np.zeros((num_hori, d.ncell))
    s = np.zeros(d.ncell)
    p = lambda x, t: x.ctypes.data_as(ctypes.POINTER(t))  # noqa: E731
    rc = L.horizon_svf_comp_ex(*(p(x, ctypes.c_double) for x in a[:5]), p(a[5], ctypes.c_int),
                               p(h, ctypes.c_double), p(s, ctypes.c_double), d.ncell, d.nvert,
                               num_hori, 1, 40000.0, 0.2, rf, 2, None, None)
    return rc, h, s


def test_c_abi_invalid_arguments_fill_nan():
    d = synthetic.lattice_domain(10.0, 10.0, 2.0)
    rc, h, s = _raw_call(d, num_hori=3)
    assert rc == _lib.HZ_EINVAL
    assert np.all(np.isnan(h)) and np.all(np.isnan(s))
    assert "num_hori" in _lib.last_error()


@pytest.mark.skipif(_has_device(), reason="checks the no-device error path")
def test_c_abi_no_device_fails_loudly():
    d = synthetic.lattice_domain(10.0, 10.0, 2.0)
    rc, h, s = _raw_call(d)
    assert rc == _lib.HZ_EDEVICE
    assert np.all(np.isnan(h)) and np.all(np.isnan(s))
    with pytest.raises(_lib.HzError):
        horizon_svf_comp_py(*_args(d))


def _fortran_call(d, num_hori=24, buf_len=4096):
    L = _lib.lib()
    h = np.empty((num_hori, d.ncell))
    s = np.empty(d.ncell)
    buf = ctypes.create_string_buffer(buf_len)
    n = ctypes.c_int(buf_len)
    p = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))  # noqa: E731
    cov = np.ascontiguousarray(d.cells_of_vertex, np.int32)
    L.horizon_svf_comp_fortran(p(d.clon, ctypes.c_double), p(d.clat, ctypes.c_double),
                               p(d.hsurf, ctypes.c_double), p(d.vlon, ctypes.c_double),
                               p(d.vlat, ctypes.c_double), p(cov, ctypes.c_int),
                               p(h, ctypes.c_double), p(s, ctypes.c_double), d.ncell, d.nvert,
                               num_hori, 1, 40000.0, 0.2, 1, 2, buf, ctypes.byref(n))
    return h, s, buf.raw[:n.value].decode()


def test_fortran_variant_reports_into_buffer():
    # reference :561-574 / :778-786: the report goes to the caller's buffer,
    # truncated to buffer_len and not NUL-terminated
    d = synthetic.lattice_domain(10.0, 10.0, 2.0)
    h, s, text = _fortran_call(d, num_hori=3)
    assert np.all(np.isnan(h)) and np.all(np.isnan(s))
    assert "failed" in text and "num_hori" in text
    _, _, short = _fortran_call(d, num_hori=3, buf_len=20)
    assert len(short) == 20 and text.startswith(short)


def test_calibrate_gather_checks_arguments():
    # the checks come before any device work, so they hold without a GPU
    from horayzon_extpar_amd._lib import calibrate_gather
    for args in ((100, 10, 10), (1 << 16, 1 << 17, 10), (1 << 16, 0, 10), (1 << 16, 16, 0),
                 (1 << 28, 16, 10)):
        with pytest.raises(_lib.HzError) as e:
            calibrate_gather(*args)
        assert e.value.code == -1
