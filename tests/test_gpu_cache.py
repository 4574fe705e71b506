"""GPU: the one-shot calls' device-buffer cache (option cache).

A call keeps its device buffers for the next call on the device whose grid
has the same sizes.  A kept domain loaded with another terrain of the same
sizes must give exactly what a fresh call gives, the certified lattice
search included (its boundary index is rebuilt).  A call of other sizes
replaces the kept domain, and hz_cache_release frees it.
"""

import numpy as np
import pytest

from horayzon_extpar_amd import _lib, synthetic
from horayzon_extpar_amd.horizon_svf import horizon_svf_comp_py

pytestmark = pytest.mark.gpu


def _run(dom, cache, **kw):
    st = {}
    h, s = horizon_svf_comp_py(*dom.args(), 24, 1, 40000.0, 0.2, 1, 2, cache=cache, stats=st,
                               **kw)
    return h, s, st


@pytest.mark.parametrize("mode", [0, 1])
def test_kept_buffers_give_fresh_results(mode):
    print(f"case cache search_mode={mode}", flush=True)
    _lib.lib().hz_cache_release()
    a = synthetic.config1(seed=11)
    b = synthetic.config1(seed=12)  # same lattice: same sizes, other terrain
    c = synthetic.config1(seed=13, size_km=60.0)  # other sizes
    assert a.ncell == b.ncell and not np.array_equal(a.hsurf, b.hsurf)
    ref = {k: _run(d, 0, search_mode=mode)[:2] for k, d in (("a", a), ("b", b), ("c", c))}
    for k, d in (("a", a), ("b", b), ("b", b), ("c", c), ("a", a), ("b", b)):
        h, s, st = _run(d, 1, search_mode=mode)
        assert np.array_equal(h, ref[k][0]) and np.array_equal(s, ref[k][1]), k
    _lib.lib().hz_cache_release()
    h, s, _ = _run(a, 1, search_mode=mode)
    assert np.array_equal(h, ref["a"][0]) and np.array_equal(s, ref["a"][1])
    _lib.lib().hz_cache_release()


def test_cache_option_checked():
    print("case cache option", flush=True)
    a = synthetic.config1(seed=11)
    with pytest.raises(_lib.HzError):
        _run(a, 2)


def test_kept_buffers_c2(c2):
    # the reload path (tree built while the host works, cell frames built on
    # the device from the host's FP64 ENU points) on the full C2 grid: the
    # same bits as a call that keeps nothing
    print("case cache c2", flush=True)
    _lib.lib().hz_cache_release()
    h0, s0, _ = _run(c2, 0)
    _run(c2, 1)
    h1, s1, st = _run(c2, 1)  # reuses the kept domain
    _lib.lib().hz_cache_release()
    assert np.array_equal(h0, h1) and np.array_equal(s0, s1)
    print(f"c2 kept: prologue {st['t_prologue_ms']:.2f} ms, upload {st['t_upload_ms']:.2f} ms, "
          f"bvh {st['t_bvh_ms']:.2f} ms", flush=True)
