"""GPU parity: libhorayzon.so (through its C-ABI) against the CPU oracle.

The oracle restates the reference walk (mo_lradtopo_horayzon.cpp:490-555)
with the same FP64 direction arithmetic and the same FP32 watertight test,
so the bar is bit-exact horizon values; the SVF differs only through the
device's sin() (tolerance 1e-12).
"""

import numpy as np
import pytest

import oracle
from horayzon_extpar_amd import synthetic
from horayzon_extpar_amd.horizon_svf import ResidentDomain, horizon_svf_comp_py

pytestmark = pytest.mark.gpu

SVF_TOL = 1e-12


def _report(name, h_gpu, h_orc):
    diff = h_gpu != h_orc
    print(f"[{name}] angles={h_gpu.size} differing={int(diff.sum())} "
          f"max|dh|={float(np.abs(h_gpu - h_orc).max()):.3g} deg", flush=True)


@pytest.mark.parametrize("grid_type", [1, 0])
@pytest.mark.parametrize("svf_type", [2, 1, 0])
def test_c1_walk_bit_exact(c1, grid_type, svf_type):
    print(f"case c1 walk gt={grid_type} svf={svf_type}", flush=True)
    st = {}
    h, s = horizon_svf_comp_py(*c1.args(), 24, grid_type, 40000.0, 0.2, 1, svf_type,
                               stats=st)
    ho, so, sto = oracle.horizon_svf(c1, grid_type=grid_type, svf_type=svf_type)
    _report(f"c1 gt{grid_type}", h, ho)
    assert np.array_equal(h, ho)
    assert np.max(np.abs(s - so)) <= SVF_TOL
    assert st["num_rays"] == sto.num_rays


def test_c1_refine_factor_10(c1):
    print("case c1 rf=10 svf=1", flush=True)
    h, s = horizon_svf_comp_py(*c1.args(), 24, 1, 40000.0, 0.2, 10, 1)
    ho, so, _ = oracle.horizon_svf(c1, grid_type=1, refine_factor=10, svf_type=1)
    _report("c1 rf10", h, ho)
    assert np.array_equal(h, ho)
    assert np.max(np.abs(s - so)) <= SVF_TOL


# the plain climb (no certificate) runs one work item per (cell, azimuth)
# and sums the sector means and the SVF in a separate kernel: sub-azimuth
# sectors and every SVF type against the oracle
@pytest.mark.parametrize("num_hori,rf,svf_type", [(24, 1, 2), (12, 3, 1), (8, 5, 0), (40, 1, 2)])
def test_c1_climb_mode(c1, num_hori, rf, svf_type):
    print(f"case c1 climb nhori={num_hori} rf={rf} svf={svf_type}", flush=True)
    st = {}
    h, s = horizon_svf_comp_py(*c1.args(), num_hori, 1, 40000.0, 0.2, rf, svf_type,
                               search_mode=1, certify=0, stats=st)
    ho, so, sto = oracle.horizon_svf(c1, num_hori=num_hori, grid_type=1, refine_factor=rf,
                                     svf_type=svf_type, mode=1)
    _report("c1 climb", h, ho)
    assert np.array_equal(h, ho)
    assert np.max(np.abs(s - so)) <= SVF_TOL
    assert st["num_rays"] == sto.num_rays


@pytest.mark.parametrize("kernel,block,refill",
                         [(1, 128, 0), (1, 64, 0), (0, 64, 64), (0, 256, 8), (0, 128, 1)])
def test_kernel_variants_bit_exact(c1, kernel, block, refill):
    print(f"case c1 kernel={kernel} block={block} refill={refill}", flush=True)
    st = {}
    h, s = horizon_svf_comp_py(*c1.args(), 24, 1, 40000.0, 0.2, 1, 2, trace_kernel=kernel,
                               block_size=block, refill=refill or None, stats=st)
    ho, so, sto = oracle.horizon_svf(c1, grid_type=1)
    _report(f"c1 kernel{kernel}", h, ho)
    assert np.array_equal(h, ho)
    assert np.max(np.abs(s - so)) <= SVF_TOL
    assert st["num_rays"] == sto.num_rays


@pytest.mark.parametrize("leaf,morton,width",
                         [(2, 0, 4), (8, 0, 4), (1, 1, 4), (4, 1, 4), (1, 0, 2), (2, 0, 2),
                          (8, 1, 2)])
def test_tree_independent(c1, leaf, morton, width):
    # the occlusion answer is the OR over all triangles: any tree gives the same bits
    print(f"case c1 leaf={leaf} morton={morton} width={width}", flush=True)
    h4, s4 = horizon_svf_comp_py(*c1.args(), 24, 1, 40000.0, 0.2, 1, 2)
    h, s = horizon_svf_comp_py(*c1.args(), 24, 1, 40000.0, 0.2, 1, 2, leaf_size=leaf,
                               block_size=64, bvh_morton=morton, bvh_width=width)
    assert np.array_equal(h, h4) and np.array_equal(s, s4)


def test_c2_binary_and_four_wide_trees_agree(c2):
    # full-size check of the quantized four-wide boxes: every one of the 13M
    # angles and the query count equal those of the full-precision binary tree
    print("case c2 width 2 vs 4", flush=True)
    with ResidentDomain(*c2.args(), grid_type=1, ray_org_elev=0.2) as dom:
        st4 = dom.run(24, 40000.0, 1, 2, bvh_width=4, collect_counters=1)
        h4, s4 = dom.fetch()
        st2 = dom.run(24, 40000.0, 1, 2, bvh_width=2, collect_counters=1)
        h2, s2 = dom.fetch()
    print(f"c2 nodes/ray: binary {st2['node_visits'] / st2['num_rays']:.1f} "
          f"four-wide {st4['node_visits'] / st4['num_rays']:.1f}; tris/ray "
          f"{st2['tri_tests'] / st2['num_rays']:.1f} / {st4['tri_tests'] / st4['num_rays']:.1f}; "
          f"nodes {st2['bvh_nodes']} / {st4['bvh_nodes']}", flush=True)
    assert np.array_equal(h4, h2) and np.array_equal(s4, s2)
    assert st4["num_rays"] == st2["num_rays"]
    assert st4["kernel_error"] == 0 and st2["kernel_error"] == 0
    assert 0 < st4["bvh_nodes"] < st2["bvh_nodes"]


def test_flat_plain():
    print("case flat plain", flush=True)
    d = synthetic.lattice_domain(40.0, 40.0, 2.0)
    h, s = horizon_svf_comp_py(*d.args(), 24, 1, 40000.0, 0.2, 1, 2)
    assert np.all(h == 0.0) and np.all(s == 1.0)


def test_c2_sampled_and_invariants(c2):
    print("case c2 sampled", flush=True)
    with ResidentDomain(*c2.args(), grid_type=1, ray_org_elev=0.2) as dom:
        st = dom.run(24, 40000.0, 1, 2)
        h, s = dom.fetch()
        st2 = dom.run(24, 40000.0, 1, 2)
        h2, s2 = dom.fetch()
    print(f"c2 rays={st['num_rays']} trace_ms={st['t_trace_ms']:.2f} "
          f"bvh_ms={st['t_bvh_ms']:.2f}", flush=True)
    assert np.array_equal(h, h2) and np.array_equal(s, s2)  # deterministic
    assert st["num_rays"] == st2["num_rays"]
    # lattice invariant at rf = 1: every value is 0 or 0.25 + 0.5 k
    k = (h - 0.25) / 0.5
    on_lattice = (h == 0.0) | (np.abs(k - np.round(k)) < 1e-9)
    assert on_lattice.all()
    assert np.all((s > 0.0) & (s <= 1.0))
    # oracle on three contiguous cell ranges
    ho = np.full_like(h, np.nan)
    so = np.full_like(s, np.nan)
    n = c2.ncell
    for b in (0, n // 2, n - 700):
        oracle.horizon_svf(c2, grid_type=1, cell_begin=b, cell_end=b + 700, out=(ho, so))
    m = ~np.isnan(so)
    _report("c2 sampled", h[:, m], ho[:, m])
    assert np.array_equal(h[:, m], ho[:, m])
    assert np.max(np.abs(s[m] - so[m])) <= SVF_TOL


def test_fortran_variant_matches(c1):
    # the EXTPAR-interface variant computes the same arrays and reports the
    # reference's ray statistics into the caller's buffer
    import ctypes
    from horayzon_extpar_amd import _lib
    print("case c1 fortran variant", flush=True)
    st = {}
    h0, s0 = horizon_svf_comp_py(*c1.args(), 24, 1, 40000.0, 0.2, 1, 2, stats=st)
    h = np.empty_like(h0)
    s = np.empty_like(s0)
    buf = ctypes.create_string_buffer(4096)
    n = ctypes.c_int(4096)
    p = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))  # noqa: E731
    cov = np.ascontiguousarray(c1.cells_of_vertex, np.int32)
    _lib.lib().horizon_svf_comp_fortran(
        p(c1.clon, ctypes.c_double), p(c1.clat, ctypes.c_double), p(c1.hsurf, ctypes.c_double),
        p(c1.vlon, ctypes.c_double), p(c1.vlat, ctypes.c_double), p(cov, ctypes.c_int),
        p(h, ctypes.c_double), p(s, ctypes.c_double), c1.ncell, c1.nvert, 24, 1, 40000.0, 0.2,
        1, 2, buf, ctypes.byref(n))
    text = buf.raw[:n.value].decode()
    print(text, flush=True)
    assert np.array_equal(h, h0) and np.array_equal(s, s0)
    assert f"Number of rays shot: {st['num_rays']}" in text
    assert f"Number of triangles in mesh: {st['num_triangles']}" in text
