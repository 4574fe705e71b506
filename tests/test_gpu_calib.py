"""GPU: the random-gather calibration behind the walk's roofline (bench.py
`roofline.gather_ceiling`): dependent random 64-byte reads of 128-byte
records at the trace kernel's occupancy give a rate of tens of G steps/s,
higher for a set that stays in the caches than for one spread over an
array larger than L2."""

import pytest

from horayzon_extpar_amd._lib import calibrate_gather

pytestmark = pytest.mark.gpu


def test_calibrate_gather_rates():
    print("case calibrate gather: 64k records hot in 16k, 4M records random", flush=True)
    hot = calibrate_gather(1 << 16, 1 << 14, iters=400)
    spread = calibrate_gather(1 << 22, 1 << 22, iters=400)
    print(f"resident set {hot / 1e9:.1f} G/s, 512 MB spread {spread / 1e9:.1f} G/s", flush=True)
    assert spread > 5e9
    assert hot > spread
