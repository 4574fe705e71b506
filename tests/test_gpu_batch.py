"""GPU: batched domains (one persistent launch, HIP graph replay, autotune).

Per-domain results of a batched run must equal separate single-domain runs
bit for bit (the occlusion answers do not depend on the launch), and every
replay of the captured graph must equal the eager batched result, with eager
work on the device between the replays.
"""

import numpy as np
import pytest

from horayzon_extpar_amd import batch as hb
from horayzon_extpar_amd import synthetic
from horayzon_extpar_amd.horizon_svf import ResidentDomain

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_c4():
    items = synthetic.config4(n_domains=8, seed=77, cells_min=2e3, cells_max=2e4)
    res = [(ResidentDomain(*d.args(), grid_type=1), nh) for d, nh in items]
    yield res
    for d, _ in res:
        d.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_batched_equals_separate(small_c4, mode):
    print(f"case batch mode={mode}", flush=True)
    groups = {}
    for d, nh in small_c4:
        groups.setdefault(nh, []).append(d)
    for nh, doms in groups.items():
        sep = []
        rays_sep = 0
        for d in doms:
            st = d.run(nh, 40000.0, 1, 2, search_mode=mode, certify=0)
            rays_sep += st["num_rays"]
            sep.append(d.fetch())
        with hb.Batch(doms) as b:
            st = b.run(nh, 40000.0, 1, 2, use_graph=False, search_mode=mode, certify=0)
            got = b.fetch()
        assert st["num_rays"] == rays_sep
        for (h0, s0), (h1, s1) in zip(sep, got):
            assert np.array_equal(h0, h1) and np.array_equal(s0, s1)


def test_graph_replays(small_c4):
    print("case batch graph replays", flush=True)
    doms = [d for d, nh in small_c4 if nh == small_c4[0][1]]
    other = [d for d, nh in small_c4 if nh != small_c4[0][1]] or doms[:1]
    nh = small_c4[0][1]
    with hb.Batch(doms) as b:
        b.run(nh, 40000.0, 1, 2, use_graph=False)
        ref = b.fetch()
        for rep in range(4):
            st = b.run(nh, 40000.0, 1, 2, use_graph=True, chunk=32 if rep < 2 else 64)
            got = b.fetch()
            for (h0, s0), (h1, s1) in zip(ref, got):
                assert np.array_equal(h0, h1) and np.array_equal(s0, s1)
            other[0].run(24, 40000.0, 1, 2)  # eager work between replays
            assert st["num_rays"] > 0


def test_autotune(small_c4):
    print("case batch autotune", flush=True)
    doms = [d for d, nh in small_c4 if nh == small_c4[0][1]]
    nh = small_c4[0][1]
    with hb.Batch(doms) as b:
        best, ms, table = hb.autotune(b, nh, space=dict(leaf_size=(1, 2), refill=(40,),
                                                          chunk=(32, 64)))
        assert len(table) == 4 and ms > 0
        assert best in [c for c, _ in table]
        b.run(nh, 40000.0, 1, 2, **best)
        got = b.fetch()
    for d, (h, s) in zip(doms, got):
        d.run(nh, 40000.0, 1, 2)
        h0, s0 = d.fetch()
        assert np.array_equal(h, h0) and np.array_equal(s, s0)
