"""Host-side parts of bench.py (no GPU): the one-GPU guard, the traffic
figure read from the committed profile, and the accuracy leg, exercised with
the oracle's own horizons standing in for the device's (so every comparison
must come out exact) and with one planted difference."""

import os
import subprocess
import sys

import numpy as np
import pytest

import bench
import oracle
from horayzon_extpar_amd import synthetic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_profiled_traffic_default_only():
    import argparse
    a = argparse.Namespace(config="c2", mode="walk", kernel=0, width=0, leaf=0, block=0,
                           morton=0)
    t = bench.profiled_traffic(a)
    assert t is not None and t[0] > 1e11
    a.mode = "climb"
    assert bench.profiled_traffic(a) is None


@pytest.fixture(scope="module")
def c1_oracle():
    d = synthetic.config1()
    h, s, _ = oracle.horizon_svf(d, grid_type=1)
    return d, h, s


def test_accuracy_leg_exact(c1_oracle):
    d, h, s = c1_oracle
    acc = bench.accuracy_leg(d, h.copy(), s.copy(), h, s, 8, nsample=12)
    assert acc["oracle_cells_compared"] == d.ncell
    assert acc["oracle_angles_differing"] == 0 and acc["oracle_max_abs_dh_deg"] == 0.0
    bf = acc["fp64_brute_force"]
    assert bf["clear_angles"] > 0
    assert bf["clear_angles_agreeing"] == bf["clear_angles"]
    assert bf["max_abs_err_vs_true_deg"] <= 0.25 + 1e-9


def test_accuracy_leg_sees_a_planted_difference(c1_oracle):
    d, h, s = c1_oracle
    hp = h.copy()
    hp[3, 7] += 0.5
    acc = bench.accuracy_leg(d, hp, s.copy(), h, s, 8, nsample=4)
    assert acc["oracle_angles_differing"] == 1
    assert abs(acc["oracle_max_abs_dh_deg"] - 0.5) < 1e-12
