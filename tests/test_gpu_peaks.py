"""pgamd_measure_peaks (needs an MI355X): the LDS and HBM bandwidths that the roofline report divides by."""
import math

import pytest

import gpu_testbed_poisson_gravity_amd as pg

pytestmark = pytest.mark.gpu


def test_measured_peaks_are_plausible():
    with pg.Context(0) as ctx:
        print("[case] measure_peaks", flush=True)
        lds, hbm = ctx.measure_peaks()
    print(f"  LDS read {lds / 1e12:.1f} TB/s, HBM copy {hbm / 1e12:.2f} TB/s", flush=True)
    assert math.isfinite(lds) and math.isfinite(hbm)
    # 256 CUs read at least 64 B/clk each at 1 GHz; no CU reads more than 512 B/clk at 2.4 GHz
    assert 256 * 64 * 1e9 < lds < 256 * 512 * 2.4e9
    # an HBM3E device copy moves more than 1 TB/s and less than the 8 TB/s interface peak
    assert 1e12 < hbm < 8e12
