"""GPU: the certified lattice searches reproduce the reference walk.

The lattice searches (climb from 0 deg, learned start) find each azimuth's
top hit on the 0.5 deg lattice; the reference walk (mo_lradtopo_horayzon.cpp
:521-546) finds the same value wherever the hits it explores are monotone in
elevation.  Along a vertical half-plane through a closed terrain surface
they are, unless the profile crosses a mesh-boundary edge (the domain border
or the holes grid_type 1 leaves at vertices with fewer than six cells,
:299).  With certify=1 (the default) the library marks every cell with a
boundary point within reach of its rays (dist_search plus the longest edge)
as near; near cells run the reference walk itself, bit for bit, and the far
cells run the lattice search, whose values then equal the walk's to its FP64
rounding of the carried elevation.

* The device equals the oracle's certified modes (3: climb, 4: learned
  start) bit for bit: horizons, rays cast and cells walked; both equal the
  reference walk to FP64 rounding (<= 1e-9 deg).
* On the full 600 x 400 km C2 terrain with the committed predictor, the
  certified learned search and the certified climb equal the walk on all
  13.3M angles, while the plain lattice search does not.
"""

import numpy as np
import pytest

import oracle
from horayzon_extpar_amd import learned, synthetic
from horayzon_extpar_amd._lib import HzError
from horayzon_extpar_amd.horizon_svf import ResidentDomain, horizon_svf_comp_py

pytestmark = pytest.mark.gpu

DH_TOL = 1e-9  # deg: the lattice value against the walk's accumulated elevation
SVF_TOL = 1e-12


@pytest.mark.parametrize("num_hori,rf,svf_type", [(24, 1, 2), (12, 3, 1), (8, 5, 0)])
def test_c1_certified_climb_equals_oracle_and_walk(c1, num_hori, rf, svf_type):
    print(f"case c1 certified climb nhori={num_hori} rf={rf}", flush=True)
    st = {}
    h, s = horizon_svf_comp_py(*c1.args(), num_hori, 1, 40000.0, 0.2, rf, svf_type,
                               search_mode=1, stats=st)
    ho, so, sto = oracle.horizon_svf(c1, num_hori=num_hori, grid_type=1, refine_factor=rf,
                                     svf_type=svf_type, mode=3)
    assert np.array_equal(h, ho)
    assert np.max(np.abs(s - so)) <= SVF_TOL
    assert st["num_rays"] == sto.num_rays
    assert int(st["cert_walk_cells"]) == sto.cells_walked
    hw, sw, _ = oracle.horizon_svf(c1, num_hori=num_hori, grid_type=1, refine_factor=rf,
                                   svf_type=svf_type, mode=0)
    print(f"walked {sto.cells_walked} of {c1.ncell} cells; max |dh| vs walk "
          f"{np.abs(h - hw).max():.3g}", flush=True)
    assert np.abs(h - hw).max() <= DH_TOL
    assert np.max(np.abs(s - sw)) <= SVF_TOL
    # the plain climb differs from the walk at the mesh border on this domain
    hc, _, _ = oracle.horizon_svf(c1, num_hori=num_hori, grid_type=1, refine_factor=rf,
                                  svf_type=svf_type, mode=1)
    assert np.abs(hc - hw).max() > 0.25


# The lattice kernels (work items and whole cells) and the certified searches'
# near-cell walk step with the next node's boxes requested early into a node
# buffer per wavefront: every block size, refill thresholds, both schedules
# (the certified search takes work items only).
@pytest.mark.parametrize("block,whole,refill", [(64, 0, 0), (256, 0, 8), (64, 1, 64), (256, 1, 0)])
def test_c1_lattice_block_sizes_equal_oracle(c1, block, whole, refill):
    print(f"case c1 climb block={block} whole_cells={whole} refill={refill}", flush=True)
    st = {}
    h, s = horizon_svf_comp_py(*c1.args(), 24, 1, 40000.0, 0.2, 1, 2, search_mode=1, certify=0,
                               block_size=block, whole_cells=whole, refill=refill or None,
                               stats=st)
    ho, so, sto = oracle.horizon_svf(c1, grid_type=1, mode=1)
    assert np.array_equal(h, ho)
    assert np.max(np.abs(s - so)) <= SVF_TOL
    assert st["num_rays"] == sto.num_rays
    if whole:
        return
    st = {}
    h, s = horizon_svf_comp_py(*c1.args(), 24, 1, 40000.0, 0.2, 1, 2, search_mode=1,
                               block_size=block, refill=refill or None, stats=st)
    ho, so, sto = oracle.horizon_svf(c1, grid_type=1, mode=3)
    assert np.array_equal(h, ho)
    assert np.max(np.abs(s - so)) <= SVF_TOL
    assert st["num_rays"] == sto.num_rays
    assert int(st["cert_walk_cells"]) == sto.cells_walked


def test_certified_learned_small_c2_equals_oracle():
    print("case small c2 certified learned", flush=True)
    dom = synthetic.config2(seed=21, width_km=120.0, height_km=90.0)
    pred = learned.Predictor.from_file(learned.DEFAULT_WEIGHTS)
    with ResidentDomain(*dom.args(), grid_type=1) as rd:
        rd.set_predictor(pred)
        st = rd.run(24, 40000.0, 1, 2, search_mode=2)
        h, s = rd.fetch()
        P = rd.predictions()
        st0 = rd.run(24, 40000.0, 1, 2, search_mode=0)
        h0, s0 = rd.fetch()
    ho, so, sto = oracle.horizon_svf(dom, grid_type=1, mode=4, pred=P)
    assert np.array_equal(h, ho) and np.max(np.abs(s - so)) <= SVF_TOL
    assert st["num_rays"] == sto.num_rays
    assert int(st["cert_walk_cells"]) == sto.cells_walked
    assert np.abs(h - h0).max() <= DH_TOL and np.max(np.abs(s - s0)) <= SVF_TOL
    print(f"rays: walk {st0['num_rays']} certified learned {st['num_rays']}; walked cells "
          f"{int(st['cert_walk_cells'])} of {dom.ncell}", flush=True)


def test_c2_full_learned_and_climb_equal_walk(c2):
    # the full C5 workload: the committed predictor on the 600 x 400 km C2
    # terrain; every one of the 13.3M angles within 1e-9 deg of the walk
    print("case c2 full certified learned / climb vs walk", flush=True)
    pred = learned.Predictor.from_file(learned.DEFAULT_WEIGHTS)
    with ResidentDomain(*c2.args(), grid_type=1, ray_org_elev=0.2) as rd:
        rd.set_predictor(pred)
        stw = rd.run(24, 40000.0, 1, 2, search_mode=0)
        hw, sw = rd.fetch()
        out = {}
        for name, mode, cert in (("learned", 2, 1), ("climb", 1, 1), ("learned plain", 2, 0)):
            st = rd.run(24, 40000.0, 1, 2, search_mode=mode, certify=cert)
            h, s = rd.fetch()
            dh = np.abs(h - hw)
            out[name] = (st, dh, np.abs(s - sw).max())
            print(f"{name}: trace {st['t_trace_ms']:.2f} ms (certificate "
                  f"{st['t_certify_ms']:.2f}), rays {st['num_rays']}, walked cells "
                  f"{int(st['cert_walk_cells'])}, angles > 1e-9 deg: {int((dh > DH_TOL).sum())}, "
                  f"max |dh| {dh.max():.3g}, max |dsvf| {out[name][2]:.3g}", flush=True)
    print(f"walk: trace {stw['t_trace_ms']:.2f} ms, rays {stw['num_rays']}", flush=True)
    for name in ("learned", "climb"):
        st, dh, dsvf = out[name]
        assert int((dh > DH_TOL).sum()) == 0, name
        assert dsvf <= SVF_TOL, name
        assert st["kernel_error"] == 0
    # without the certificate the lattice search misses the walk at the border
    assert int((out["learned plain"][1] > DH_TOL).sum()) > 0
    assert out["learned"][0]["num_rays"] < stw["num_rays"]


def test_certify_option_checks(c1):
    print("case certify option checks", flush=True)
    with ResidentDomain(*c1.args(), grid_type=1) as rd:
        with pytest.raises(HzError):
            rd.run(24, 40000.0, 1, 2, search_mode=1, whole_cells=1)
        with pytest.raises(HzError):
            rd.run(24, 40000.0, 1, 2, search_mode=1, certify=2)
        with pytest.raises(HzError):
            rd.run(24, 40000.0, 1, 2, search_mode=1, trace_kernel=1)
        # plain lattice searches still take whole cells and the lane-per-cell kernel
        st = rd.run(24, 40000.0, 1, 2, search_mode=1, whole_cells=1, certify=0)
        assert st["cert_walk_cells"] == 0
        # the walk ignores the option
        rd.run(24, 40000.0, 1, 2, search_mode=0, certify=1)
