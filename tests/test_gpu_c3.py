"""GPU parity on the fine-mesh stress workload (config 3): 100 m core plus a
500 m outer ring joined at a seam, steep field with cliffs, nhori 48.

* bit-exact against the CPU oracle on cell blocks in the core, at the seam
  and in the ring;
* the FP64 "true horizon" (half-plane sweep over the triangles) of 2,000 random
  core cells, quantised to the 0.5 deg lattice, equals the device walk on
  every angle that is not within 1e-3 deg of a lattice elevation nor near
  the seam's mesh holes (grazing and border angles counted apart);
* lattice invariant and run-to-run determinism on all 977k cells;
* the device prologue gives the host prologue's FP32 scene and horizons.
"""

import numpy as np
import pytest

import oracle
from horayzon_extpar_amd import synthetic
from horayzon_extpar_amd.horizon_svf import ResidentDomain

pytestmark = pytest.mark.gpu
NHORI = 48


@pytest.fixture(scope="module")
def c3_run():
    d = synthetic.config3()
    with ResidentDomain(*d.args(), grid_type=1) as rd:
        st = rd.run(NHORI, 40000.0, 1, 2)
        h, s = rd.fetch()
        st2 = rd.run(NHORI, 40000.0, 1, 2)
        h2, s2 = rd.fetch()
    print(f"c3: {d.ncell} cells, rays {st['num_rays']}, trace {st['t_trace_ms']:.1f} ms, "
          f"bvh {st['t_bvh_ms']:.2f} ms", flush=True)
    assert np.array_equal(h, h2) and np.array_equal(s, s2)
    return d, h, s


def test_c3_device_prologue(c3_run):
    # the device prologue on the seam / hole-rich mesh: same FP32 scene and
    # ray origins as the host prologue, frames to 1e-15, same horizons
    from horayzon_extpar_amd.horizon_svf import scene_export
    print("case c3 device prologue", flush=True)
    d, h, s = c3_run
    host = scene_export(*d.args(), grid_type=1)
    with ResidentDomain(*d.args(), grid_type=1, prologue="device") as rd:
        dev = rd.scene()
        rd.run(NHORI, 40000.0, 1, 2)
        hd, sd = rd.fetch()
    assert np.array_equal(dev["tri"], host["tri"])
    assert np.array_equal(dev["vert32"].view(np.int32), host["vert32"].view(np.int32))
    assert np.array_equal(dev["org"].view(np.int32), host["org"].view(np.int32))
    assert np.max(np.abs(dev["frame_n"] - host["frame_n"])) <= 1e-15
    assert np.max(np.abs(dev["frame_north"] - host["frame_north"])) <= 1e-15
    assert np.array_equal(hd, h) and np.max(np.abs(sd - s)) <= 1e-12


def test_c3_lattice(c3_run):
    d, h, s = c3_run
    k = (h - 0.25) / 0.5
    assert np.all((h == 0.0) | (np.abs(k - np.round(k)) < 1e-9))
    assert np.all((s > 0.0) & (s <= 1.0))


def test_c3_oracle_blocks(c3_run):
    print("case c3 oracle blocks", flush=True)
    d, h, s = c3_run
    ho = np.full_like(h, np.nan)
    so = np.full_like(s, np.nan)
    ncore = d.meta["ncore"]
    starts = [ncore // 2, ncore - 400, ncore + 20000]  # core centre, seam, ring
    for b in starts:
        oracle.horizon_svf(d, num_hori=NHORI, grid_type=1, cell_begin=b, cell_end=b + 400,
                           out=(ho, so))
    m = ~np.isnan(so)
    assert m.sum() == 1200
    assert np.array_equal(h[:, m], ho[:, m])
    assert np.max(np.abs(s[m] - so[m])) <= 1e-12


def test_c3_fp64_true_horizon(c3_run):
    # the FP64 half-plane horizon of 2,000 random cells of the inner core
    # (their 40 km search never reaches the outer edge of the ring), quantised
    # to the walk's 0.5 deg lattice: the walk returns that value on every
    # angle that is neither grazing (within 1e-3 deg of a lattice elevation)
    # nor on a search segment passing within two ring edges (1 km) of a
    # mesh-boundary vertex (the seam between core and ring), where the walk can
    # see through a hole and its result depends on its path
    from horayzon_extpar_amd import diagnostics
    d, h, s = c3_run
    rng = np.random.default_rng(3)
    core = np.flatnonzero((np.abs(d.cx) < 20e3) & (np.abs(d.cy) < 20e3))
    cells = np.sort(rng.choice(core, 2000, replace=False))
    print(f"case c3 fp64 true horizon on {cells.size} cells", flush=True)
    th = np.degrees(oracle.true_horizon(d, cells, num_hori=NHORI, grid_type=1))
    exp = np.where(th > 0, 0.25 + 0.5 * np.floor(th / 0.5), 0.0)
    grazing = np.abs(th / 0.5 - np.round(th / 0.5)) * 0.5 <= 1e-3
    az = np.arange(NHORI) * (360.0 / NHORI)
    border = diagnostics.border_azimuths(d.clon, d.clat, d.vlon, d.vlat, d.cells_of_vertex,
                                         cells, az, 40000.0, 1000.0) & ~grazing
    clear = ~grazing & ~border
    got = h[:, cells].T
    agree = np.abs(got - exp) < 1e-9
    print(f"c3 fp64: clear {int(agree[clear].sum())} of {int(clear.sum())} agree; grazing "
          f"{int(agree[grazing].sum())} of {int(grazing.sum())}; border "
          f"{int(agree[border].sum())} of {int(border.sum())}", flush=True)
    # most 40 km segments of the inner core cross the seam, 30 km out
    assert clear.sum() > 0.2 * clear.size
    assert agree[clear].all()
