"""Concurrent calls: the C entry points are reentrant (each call owns its
stream, buffers and staging memory; the last-error text is per thread), so
calls from several host threads at once give the serial results bit for bit.
ctypes releases the GIL during the foreign call, so the calls overlap."""

import threading

import numpy as np
import pytest

from horayzon_extpar_amd import synthetic
from horayzon_extpar_amd.horizon_svf import horizon_svf_comp_py

pytestmark = pytest.mark.gpu


def test_concurrent_calls_match_serial(c1):
    print("case 4 threads x 2 calls", flush=True)
    rng = np.random.default_rng(17)
    a = synthetic.lattice_domain(30.0, 20.0, 1.0)
    a.hsurf = rng.uniform(0.0, 800.0, a.ncell)
    b = synthetic.lattice_domain(12.0, 12.0, 0.5)
    b.hsurf = rng.uniform(0.0, 300.0, b.ncell)
    doms = [c1, a, b, synthetic.tilted_plane(30.0, 1.0, 0.05, -0.02)]
    kw = [dict(grid_type=1, num_hori=24), dict(grid_type=0, num_hori=48),
          dict(grid_type=1, num_hori=12), dict(grid_type=1, num_hori=24)]

    def call(i):
        return horizon_svf_comp_py(*doms[i].args(), kw[i]["num_hori"], kw[i]["grid_type"],
                                   40000.0, 0.2, 1, 2)

    serial = [call(i) for i in range(len(doms))]
    out = [[None, None] for _ in doms]
    errors = []

    def worker(i):
        try:
            for r in range(2):
                out[i][r] = call(i)
        except Exception as e:  # reported below, in the main thread
            errors.append((i, repr(e)))

    th = [threading.Thread(target=worker, args=(i,)) for i in range(len(doms))]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=240)
    assert not any(t.is_alive() for t in th)
    assert not errors, errors
    for i in range(len(doms)):
        for r in range(2):
            h, s = out[i][r]
            assert np.array_equal(h, serial[i][0]) and np.array_equal(s, serial[i][1]), (i, r)
