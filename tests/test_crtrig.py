"""The device prologue's correctly rounded sin / cos (csrc/crtrig.h), on the
CPU: the golden fixture tests/golden/sincos_cr.npz is pinned independently,
and the library's host evaluation of the routine must reproduce it bit for
bit (the device evaluation is checked against the same fixture in
test_gpu_prologue_dev.py).
"""

import ctypes
import decimal
import math
import os

import numpy as np
import pytest

from horayzon_extpar_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "sincos_cr.npz")


def _fixture():
    z = np.load(FIX, allow_pickle=False)
    return z["x"], z["sin"], z["cos"]


def _ulps(a, b):
    return np.abs(a.view(np.int64) - b.view(np.int64))


def _sin_triple(x: float) -> float:
    """sin x by a different route than the fixture's: Taylor series of
    x / 3^8 (no pi reduction), then the triple-angle identity
    sin 3t = 3 sin t - 4 sin^3 t eight times, at 80 digits."""
    with decimal.localcontext() as ctx:
        ctx.prec = 80
        t = decimal.Decimal(x) / (3 ** 8)
        s, term, n = t, t, 1
        while abs(term) > decimal.Decimal(10) ** -90:
            term = -term * t * t / ((2 * n) * (2 * n + 1))
            s += term
            n += 1
        for _ in range(8):
            s = 3 * s - 4 * s * s * s
        return float(s)


def test_fixture_independent_route():
    # every fixture sine of |x| <= 4 recomputed through the triple-angle route
    x, s, _ = _fixture()
    sel = np.nonzero(np.abs(x) <= 4.0)[0][::3]
    assert sel.size > 500
    for i in sel:
        assert _sin_triple(float(x[i])) == s[i], float(x[i])


def test_fixture_closed_forms():
    x, s, c = _fixture()
    for xv, sv, cv in zip(x, s, c):
        if xv == 0.0:
            assert sv == xv and math.copysign(1.0, sv) == math.copysign(1.0, xv) and cv == 1.0
    # sin^2 + cos^2 = 1 to rounding, and the fixture is within 1 ulp of glibc
    assert np.max(np.abs(s * s + c * c - 1.0)) < 4e-16
    gs = np.array([math.sin(v) for v in x])
    gc = np.array([math.cos(v) for v in x])
    assert _ulps(gs, s).max() <= 1 and _ulps(gc, c).max() <= 1
    # glibc is not correctly rounded everywhere, but nearly
    assert (gs != s).mean() < 0.01 and (gc != c).mean() < 0.01


def test_host_evaluation_matches_fixture():
    x, s, c = _fixture()
    x = np.ascontiguousarray(x)
    hs, hc = np.empty_like(x), np.empty_like(x)
    p = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert _lib.lib().hz_sincos(p(x), x.size, p(hs), p(hc), 0) == 0
    assert np.array_equal(hs.view(np.int64), s.view(np.int64))
    assert np.array_equal(hc.view(np.int64), c.view(np.int64))


def test_bad_arguments_rejected():
    assert _lib.lib().hz_sincos(None, 3, None, None, 0) == _lib.HZ_EINVAL
    assert _lib.lib().hz_sincos(None, 0, None, None, 0) == 0


@pytest.mark.parametrize("value", ["host", "device", 0, 1])
def test_prologue_option(value):
    o = _lib.default_options(prologue=value)
    assert o.prologue == {"host": 0, "device": 1}.get(value, value)
    with pytest.raises(ValueError):
        _lib.default_options(prologue="gpu")
