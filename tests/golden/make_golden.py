"""Write the golden fixtures of tests/golden/ from the CPU oracle.

The reference has no golden vectors (no tests, no data files), so these are
regression vectors of the oracle: the horizon and SVF of small seeded
synthetic domains.  The inputs are regenerated from their seeds by
horayzon_extpar_amd.synthetic; an input digest is stored to detect drift of
the generator.  Run:  python tests/golden/make_golden.py
"""

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

CASES = [
    # name, domain kwargs, call kwargs
    ("c1s_gt1_rf1_svf2", dict(size_km=40.0, seed=303), dict(grid_type=1, refine_factor=1, svf_type=2)),
    ("c1s_gt0_rf1_svf0", dict(size_km=40.0, seed=303), dict(grid_type=0, refine_factor=1, svf_type=0)),
    ("c1s_gt1_rf3_svf1", dict(size_km=40.0, seed=304), dict(grid_type=1, refine_factor=3, svf_type=1)),
]


def digest(dom) -> str:
    h = hashlib.sha256()
    for a in dom.args():
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main():
    import oracle
    from horayzon_extpar_amd import synthetic
    for name, dkw, ckw in CASES:
        dom = synthetic.config1(**dkw)
        h, s, st = oracle.horizon_svf(dom, num_hori=24, dist_search=40000.0,
                                      ray_org_elev=0.2, **ckw)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), horizon=h, skyview=s,
                            num_rays=np.int64(st.num_rays), digest=np.array(digest(dom)))
        print(name, dom.ncell, st.num_rays)


if __name__ == "__main__":
    main()
