"""GPU: the full batched-domains workload (config 4).

64 domains (4,968 .. 345,978 cells, 2 / 1 / 0.5 km edges, nhori 24 or 48,
4.53M cells in all), each resident on the device.  Per nhori group, one
persistent launch over the group, replayed from a captured HIP graph, gives
every domain the same bits as its own separate launch; and the separate
launches equal the CPU oracle on the smallest domain (all cells), on blocks
of the largest domain, and on blocks of an nhori-48 domain.
"""

import numpy as np
import pytest

import oracle
from horayzon_extpar_amd import batch as hb
from horayzon_extpar_amd import synthetic

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c4():
    items = synthetic.config4()
    res, t = hb.make_resident(items)
    sep = []
    for d, nh in res:
        d.run(nh, 40000.0, 1, 2)
        sep.append(d.fetch())
    print(f"c4: {len(items)} domains, {sum(d.ncell for d, _ in items)} cells, resident in "
          f"{t:.2f} s", flush=True)
    yield items, res, sep
    for d, _ in res:
        d.close()


def test_c4_batched_graph_equals_separate(c4):
    items, res, sep = c4
    for nh in (24, 48):
        idx = [k for k, (_, n) in enumerate(res) if n == nh]
        doms = [res[k][0] for k in idx]
        print(f"case c4 batch nhori={nh}: {len(doms)} domains", flush=True)
        with hb.Batch(doms) as b:
            for rep in range(3):  # the first run captures the graph, the others replay it
                b.run(nh, 40000.0, 1, 2, use_graph=True)
                got = b.fetch()
                for k, (h1, s1) in zip(idx, got):
                    h0, s0 = sep[k]
                    assert np.array_equal(h0, h1) and np.array_equal(s0, s1), (rep, k)


def _oracle_blocks(dom, nh, h, s, starts, n=300):
    ho = np.full_like(h, np.nan)
    so = np.full_like(s, np.nan)
    for b in starts:
        oracle.horizon_svf(dom, num_hori=nh, grid_type=1, cell_begin=b,
                           cell_end=min(b + n, dom.ncell), out=(ho, so))
    m = ~np.isnan(so)
    assert m.any()
    assert np.array_equal(h[:, m], ho[:, m])
    assert np.max(np.abs(s[m] - so[m])) <= 1e-12
    return int(m.sum())


def test_c4_oracle_smallest_largest_nhori48(c4):
    items, res, sep = c4
    sizes = [d.ncell for d, _ in items]
    small = int(np.argmin(sizes))
    large = int(np.argmax(sizes))
    n48 = [k for k, (_, nh) in enumerate(items) if nh == 48 and k not in (small, large)]
    k48 = n48[len(n48) // 2]
    for k, starts in ((small, [0]), (large, None), (k48, None)):
        dom, nh = items[k]
        if starts is None:
            starts = [0, dom.ncell // 2, dom.ncell - 300]
        n = dom.ncell if k == small else 300
        print(f"case c4 oracle domain {k}: {dom.ncell} cells, nhori {nh}", flush=True)
        done = _oracle_blocks(dom, nh, *sep[k], starts, n)
        assert done == (dom.ncell if k == small else 900)
