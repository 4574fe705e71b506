"""GPU edge cases against the oracle (bit-exact horizons, SVF to 1e-12).

* tiny domains with 0, 1 and a handful of triangles (grid_type 0 and 1): the
  empty-scene path (every query misses, the walk runs down to -85 deg and the
  clipped horizon is 0) and the one-triangle tree;
* many azimuths (nhori 72, refine_factor 2, svf_type 1);
* the tilted plane, whose horizon has a closed form in the oracle tests;
* repeated runs on one resident domain with different parameters.
"""

import numpy as np
import pytest

import oracle
from horayzon_extpar_amd import synthetic
from horayzon_extpar_amd.horizon_svf import ResidentDomain, horizon_svf_comp_py

pytestmark = pytest.mark.gpu

SVF_TOL = 1e-12


def _check(dom, grid_type=1, num_hori=24, rf=1, svf_type=2, **kw):
    st = {}
    h, s = horizon_svf_comp_py(*dom.args(), num_hori, grid_type, 40000.0, 0.2, rf, svf_type,
                               stats=st, **kw)
    ho, so, sto = oracle.horizon_svf(dom, num_hori=num_hori, grid_type=grid_type,
                                     refine_factor=rf, svf_type=svf_type)
    assert np.array_equal(h, ho)
    assert np.max(np.abs(s - so)) <= SVF_TOL
    assert st["num_rays"] == sto.num_rays
    return h, s, st


def _three_cells():
    """Three cells around one ICON vertex: grid_type 0 builds one triangle,
    grid_type 1 none (the vertex is not 6-valent)."""
    d = synthetic.lattice_domain(6.0, 6.0, 2.0, shuffle=False)
    v = int(np.argmax((d.cells_of_vertex >= 0).sum(0)))
    cells = [int(c) for c in d.cells_of_vertex[:, v] if c >= 0][:3]
    cov = np.full((6, 1), -2, np.int32)
    cov[:3, 0] = np.arange(3)
    return synthetic.Domain(clon=d.clon[cells].copy(), clat=d.clat[cells].copy(),
                            hsurf=np.array([100.0, 250.0, 40.0]), vlon=d.vlon[v:v + 1].copy(),
                            vlat=d.vlat[v:v + 1].copy(), cells_of_vertex=cov)


@pytest.mark.parametrize("grid_type", [0, 1])
def test_three_cells(grid_type):
    print(f"case three cells gt{grid_type}", flush=True)
    h, s, st = _check(_three_cells(), grid_type=grid_type)
    assert st["num_triangles"] == (1 if grid_type == 0 else 0)
    if grid_type == 1:  # empty scene: nothing occludes
        assert np.all(h == 0.0) and np.all(s == 1.0)


@pytest.mark.parametrize("size,grid_type", [(4.0, 0), (4.0, 1), (8.0, 0), (8.0, 1)])
def test_tiny_lattices(size, grid_type):
    print(f"case tiny {size} km gt{grid_type}", flush=True)
    d = synthetic.lattice_domain(size, size, 2.0)
    rng = np.random.default_rng(5)
    d.hsurf = rng.uniform(0.0, 400.0, d.ncell)
    _, _, st = _check(d, grid_type=grid_type)
    print(f"  ncell {d.ncell} triangles {st['num_triangles']}", flush=True)


def test_many_azimuths(c1):
    print("case c1 nhori 72 rf 2 svf 1", flush=True)
    _check(c1, num_hori=72, rf=2, svf_type=1)


@pytest.mark.parametrize("rf", [1, 3])
def test_tilted_plane(rf):
    print(f"case tilted plane rf {rf}", flush=True)
    d = synthetic.tilted_plane(30.0, 1.0, 0.05, -0.02)
    _check(d, rf=rf)


def test_resident_domain_parameter_changes(c1):
    # one resident domain, runs with growing and shrinking nhori and all SVF
    # forms: each equals the oracle (outputs are re-sized, state is reset)
    print("case resident reruns", flush=True)
    with ResidentDomain(*c1.args(), grid_type=1, ray_org_elev=0.2) as dom:
        for nh, rf, svf in ((12, 1, 0), (48, 1, 2), (24, 2, 1), (12, 1, 2)):
            dom.run(nh, 40000.0, rf, svf)
            h, s = dom.fetch()
            ho, so, _ = oracle.horizon_svf(c1, num_hori=nh, grid_type=1, refine_factor=rf,
                                           svf_type=svf)
            assert h.shape == (nh, c1.ncell)
            assert np.array_equal(h, ho), (nh, rf, svf)
            assert np.max(np.abs(s - so)) <= SVF_TOL
