"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle reproduces every fixture bit for bit and the generator still
produces the same inputs.  GPU: libhorayzon.so reproduces every fixture
(horizon bit-exact, SVF within 1e-12) without consulting the oracle.
"""

import os
import sys

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)

import make_golden  # noqa: E402
import oracle  # noqa: E402
from horayzon_extpar_amd import synthetic  # noqa: E402

CASES = make_golden.CASES


def _load(name):
    z = np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)
    return z["horizon"], z["skyview"], int(z["num_rays"]), str(z["digest"])


@pytest.mark.parametrize("name,dkw,ckw", CASES, ids=[c[0] for c in CASES])
def test_oracle_reproduces_golden(name, dkw, ckw):
    h0, s0, rays0, dig = _load(name)
    dom = synthetic.config1(**dkw)
    assert make_golden.digest(dom) == dig
    h, s, st = oracle.horizon_svf(dom, num_hori=24, dist_search=40000.0, ray_org_elev=0.2,
                                  **ckw)
    assert np.array_equal(h, h0) and np.array_equal(s, s0)
    assert st.num_rays == rays0


@pytest.mark.gpu
@pytest.mark.parametrize("name,dkw,ckw", CASES, ids=[c[0] for c in CASES])
def test_gpu_reproduces_golden(name, dkw, ckw):
    from horayzon_extpar_amd.horizon_svf import horizon_svf_comp_py
    print(f"case golden {name}", flush=True)
    h0, s0, rays0, _ = _load(name)
    dom = synthetic.config1(**dkw)
    st = {}
    h, s = horizon_svf_comp_py(*dom.args(), 24, ckw["grid_type"], 40000.0, 0.2,
                               ckw["refine_factor"], ckw["svf_type"], stats=st)
    assert np.array_equal(h, h0)
    assert np.max(np.abs(s - s0)) <= 1e-12
    assert st["num_rays"] == rays0
