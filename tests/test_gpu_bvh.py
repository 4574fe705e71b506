"""Structure of the device-built traversal trees (binary LBVH and the
four-wide collapse), read back with hz_domain_bvh and walked on the host:

* every triangle is reached exactly once, no node twice (and every node
  once where the tree has no collapsed leaves);
* every child box (binary: FP32 planes, z' = z + (x^2 + y^2) / 2R in the z
  planes; four-wide: the 8-bit planes decoded exactly in FP64) contains all
  vertices of the child's subtree;
* inline triangles equal the leaf-order array, and the "wide" bit on a
  reference says exactly whether the node holds inline triangles;
* a binary node with two interior children stores the leaf-order split of
  its children (-1 only in a top rebuilt with the SAH option);
* the leaf order is a permutation of the mesh's triangles, with their
  vertices (hz_scene_export) unchanged;
* all of this also holds with the top of the binary tree rebuilt by the
  SAH (option bvh_sah), which returns the same horizons, rays and SVF as
  the Morton tree on C1 and C2.
The query results do not depend on the tree (test_gpu_parity.py); this
pins the tree itself.
"""

import sys

import numpy as np
import pytest

from horayzon_extpar_amd import synthetic
from horayzon_extpar_amd.horizon_svf import ResidentDomain, scene_export

pytestmark = pytest.mark.gpu

LEAF_MASK = (1 << 28) - 1
WIDE = 1 << 30
B4_INLINE = 1 << 30
B4_SHIFT = 27
B4_MASK = (1 << 27) - 1

BIN = np.dtype([("l", "<f4", (3, 2)), ("h", "<f4", (3, 2)), ("r", "<i4", (2,)),
                ("tri", "<f4", (18,))])
# binary node memory order: lx0 lx1 hx0 hx1 ly0 ly1 hy0 hy1 lz0 lz1 hz0 hz1
BIN_RAW = np.dtype([("box", "<f4", (3, 4)), ("r", "<i4", (2,)), ("tri", "<f4", (18,))])
B4 = np.dtype([("org", "<f4", (3,)), ("meta", "<u4"), ("q", "<u4", (6,)), ("ref", "<i4", (4,)),
               ("tri", "<f4", (18,))])


def _leaf_tris(tris):
    v = tris[:, :9].astype(np.float64).reshape(-1, 3, 3)
    ids = tris[:, 9].view(np.int32)
    return v, ids


# binary trees bound z' = z + (x^2 + y^2) * KZCURV in their z planes
KZCURV = 1.0 / (2.0 * 6371229.0)


def _check_binary(nodes_raw, tris, sah=False):
    nodes = nodes_raw.view(BIN_RAW).reshape(-1)
    v, _ = _leaf_tris(tris)
    v = v.copy()
    v[:, :, 2] += (v[:, :, 0] ** 2 + v[:, :, 1] ** 2) * KZCURV
    seen = np.zeros(len(tris), np.int64)
    visits = np.zeros(len(nodes), np.int64)

    def walk(i):
        # returns the subtree's bounds, whether node i holds an inline
        # triangle, and its leaf-order range when that is contiguous (else None)
        visits[i] += 1
        nd = nodes[i]
        bmin, bmax, inline = np.full(3, np.inf), np.full(3, -np.inf), False
        ranges = []
        for s in range(2):
            lo = nd["box"][:, s].astype(np.float64)
            hi = nd["box"][:, 2 + s].astype(np.float64)
            r = int(nd["r"][s])
            if r >= 0:
                cmin, cmax, cin, crange = walk(r & LEAF_MASK)
                assert bool(r & WIDE) == cin
            else:
                code = ~r
                first, cnt = code & LEAF_MASK, (code >> 28) + 1
                seen[first:first + cnt] += 1
                if cnt == 1:
                    inline = True
                    assert np.array_equal(nd["tri"][9 * s:9 * s + 9].view(np.int32),
                                          tris[first, :9].view(np.int32))
                cmin = v[first:first + cnt].min(axis=(0, 1))
                cmax = v[first:first + cnt].max(axis=(0, 1))
                crange = (first, first + cnt - 1)
            assert np.all(lo <= cmin) and np.all(hi >= cmax), (i, s)
            bmin, bmax = np.minimum(bmin, cmin), np.maximum(bmax, cmax)
            ranges.append(crange)
        if int(nd["r"][0]) >= 0 and int(nd["r"][1]) >= 0:
            # two interior children: the leaf split that orders the traversal
            split = int(nd["tri"][1:2].view(np.int32)[0])
            if split >= 0:
                assert ranges[0] is not None and ranges[1] is not None, i
                assert ranges[0][1] == split and ranges[1][0] == split + 1, (i, split, ranges)
            else:
                assert sah and split == -1, (i, split)
        whole = None
        if ranges[0] is not None and ranges[1] is not None and ranges[0][1] + 1 == ranges[1][0]:
            whole = (ranges[0][0], ranges[1][1])
        return bmin, bmax, inline, whole

    walk(0)
    return seen, visits


def _check_four(nodes_raw, tris):
    nodes = nodes_raw.view(B4).reshape(-1)
    v, _ = _leaf_tris(tris)
    seen = np.zeros(len(tris), np.int64)
    visits = np.zeros(len(nodes), np.int64)

    def walk(i):
        visits[i] += 1
        nd = nodes[i]
        meta = int(nd["meta"])
        nch = meta >> 24
        assert 1 <= nch <= 4
        scale = [2.0 ** (((meta >> (8 * a)) & 0xFF) - 127) for a in range(3)]
        bmin, bmax, inline = np.full(3, np.inf), np.full(3, -np.inf), False
        for k in range(nch):
            lo = np.array([float(nd["org"][a]) + ((int(nd["q"][2 * a]) >> (8 * k)) & 0xFF) * scale[a]
                           for a in range(3)])
            hi = np.array([float(nd["org"][a])
                           + ((int(nd["q"][2 * a + 1]) >> (8 * k)) & 0xFF) * scale[a]
                           for a in range(3)])
            r = int(nd["ref"][k])
            if r >= 0:
                cmin, cmax, cin = walk(r & LEAF_MASK)
                assert bool(r & WIDE) == cin
            else:
                code = ~r
                if code & B4_INLINE:
                    slot, leaf = code & 1, (code >> 1) & B4_MASK
                    inline = True
                    assert np.array_equal(nd["tri"][9 * slot:9 * slot + 9].view(np.int32),
                                          tris[leaf, :9].view(np.int32))
                    first, cnt = leaf, 1
                else:
                    first, cnt = code & B4_MASK, (code >> B4_SHIFT) + 1
                seen[first:first + cnt] += 1
                cmin = v[first:first + cnt].min(axis=(0, 1))
                cmax = v[first:first + cnt].max(axis=(0, 1))
            assert np.all(lo <= cmin) and np.all(hi >= cmax), (i, k)
            bmin, bmax = np.minimum(bmin, cmin), np.maximum(bmax, cmax)
        return bmin, bmax, inline

    walk(0)
    return seen, visits


@pytest.fixture(scope="module")
def c1_dom(c1):
    sys.setrecursionlimit(10000)
    scene = scene_export(*c1.args(), grid_type=1)
    with ResidentDomain(*c1.args(), grid_type=1, ray_org_elev=0.2) as dom:
        yield dom, scene


@pytest.mark.parametrize("width,leaf,morton,sah", [(2, 1, 0, 0), (2, 4, 0, 0), (2, 1, 1, 0),
                                                   (4, 1, 0, 0), (4, 2, 0, 0), (4, 8, 1, 0),
                                                   (2, 1, 0, 1), (2, 4, 1, 1)])
def test_tree_structure(c1_dom, width, leaf, morton, sah):
    print(f"case bvh c1 width {width} leaf {leaf} morton {morton} sah {sah}", flush=True)
    dom, scene = c1_dom
    nodes, tris = dom.bvh(bvh_width=width, leaf_size=leaf, bvh_morton=morton, bvh_sah=sah)
    ntri = scene["tri"].shape[0]
    assert tris.shape[0] == ntri and nodes.shape[0] >= 1
    seen, visits = (_check_binary(nodes, tris, sah=bool(sah)) if width == 2
                    else _check_four(nodes, tris))
    assert np.all(seen == 1)
    # no node is reached twice; with one triangle per leaf every node is
    # reached (larger binary leaves leave the nodes below them unreferenced)
    assert visits.max() == 1
    if leaf == 1 or width == 4:
        assert np.all(visits == 1)
    if width == 2:
        assert nodes.shape[0] == ntri - 1
    # leaf order: a permutation of the mesh triangles, vertices unchanged
    v, ids = _leaf_tris(tris)
    assert np.array_equal(np.sort(ids), np.arange(ntri))
    ref = scene["vert32"][scene["tri"][ids]].astype(np.float64)
    assert np.array_equal(v, ref)
    print(f"  {nodes.shape[0]} nodes for {ntri} triangles", flush=True)


def test_single_triangle_tree():
    # three cells around one ICON vertex, grid_type 0: one triangle, one node
    print("case bvh one triangle", flush=True)
    d = synthetic.lattice_domain(6.0, 6.0, 2.0, shuffle=False)
    vv = int(np.argmax((d.cells_of_vertex >= 0).sum(0)))
    cells = [int(c) for c in d.cells_of_vertex[:, vv] if c >= 0][:3]
    cov = np.full((6, 1), -2, np.int32)
    cov[:3, 0] = np.arange(3)
    dd = synthetic.Domain(clon=d.clon[cells].copy(), clat=d.clat[cells].copy(),
                          hsurf=np.array([100.0, 250.0, 40.0]), vlon=d.vlon[vv:vv + 1].copy(),
                          vlat=d.vlat[vv:vv + 1].copy(), cells_of_vertex=cov)
    with ResidentDomain(*dd.args(), grid_type=0, ray_org_elev=0.2) as dom:
        for width in (2, 4):
            nodes, tris = dom.bvh(bvh_width=width)
            assert nodes.shape[0] == 1 and tris.shape[0] == 1
            seen, visits = (_check_binary if width == 2 else _check_four)(nodes, tris)
            # the binary root of a one-triangle tree holds it in both slots
            assert seen.tolist() == [2 if width == 2 else 1] and visits.tolist() == [1]


def test_sah_top_same_results(c1, c2):
    # the SAH-rebuilt top changes which nodes a ray visits, never its answer
    for name, dom in (("c1", c1), ("c2", c2)):
        with ResidentDomain(*dom.args(), grid_type=1, ray_org_elev=0.2) as rd:
            for mode in (0, 1):
                st0 = rd.run(24, 40000.0, 1, 2, search_mode=mode, certify=0, collect_counters=1)
                h0, s0 = rd.fetch()
                st1 = rd.run(24, 40000.0, 1, 2, search_mode=mode, certify=0, collect_counters=1,
                             bvh_sah=1)
                h1, s1 = rd.fetch()
                print(f"case sah {name} mode {mode}: nodes per query "
                      f"{st0['node_visits'] / st0['num_rays']:.2f} (Morton) "
                      f"{st1['node_visits'] / st1['num_rays']:.2f} (SAH top)", flush=True)
                assert np.array_equal(h0, h1) and np.array_equal(s0, s1)
                assert st0["num_rays"] == st1["num_rays"]
        if name == "c2":
            with ResidentDomain(*dom.args(), grid_type=1, ray_org_elev=0.2) as rd:
                a = rd.run(24, 40000.0, 1, 2)["t_trace_ms"]
                b = rd.run(24, 40000.0, 1, 2, bvh_sah=1)["t_trace_ms"]
                a = min(a, rd.run(24, 40000.0, 1, 2)["t_trace_ms"])
                b = min(b, rd.run(24, 40000.0, 1, 2, bvh_sah=1)["t_trace_ms"])
                print(f"c2 walk trace: {a:.2f} ms (Morton), {b:.2f} ms (SAH top)", flush=True)


def test_sah_option_checks(c1_dom):
    from horayzon_extpar_amd._lib import HzError
    dom, _ = c1_dom
    with pytest.raises(HzError):
        dom.run(24, 40000.0, 1, 2, bvh_sah=1, bvh_width=4)
    with pytest.raises(HzError):
        dom.run(24, 40000.0, 1, 2, bvh_sah=2)
