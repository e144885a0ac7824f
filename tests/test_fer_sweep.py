"""Statistics helpers of the FER studies (cusat_amd/fer_sweep.py), on the
CPU: the Wilson score interval against its closed form at known points,
and the log-linear FER crossing on a synthetic waterfall."""
import math

import pytest

from cusat_amd import fer_sweep as fs


def test_wilson_interval_against_scipy():
    from scipy.stats import binomtest
    for k, n in ((50, 1000), (0, 2_000_000), (1000, 1_000_000), (3, 10), (10, 10)):
        ref = binomtest(k, n).proportion_ci(confidence_level=0.95, method="wilson")
        lo, hi = fs.wilson(k, n)
        assert lo == pytest.approx(ref.low, rel=1e-6, abs=1e-12), (k, n)
        assert hi == pytest.approx(ref.high, rel=1e-6), (k, n)
    # large n: close to the normal approximation
    lo, hi = fs.wilson(1000, 1_000_000)
    assert hi - lo == pytest.approx(2 * 1.96 * math.sqrt(1e-3 * (1 - 1e-3) / 1e6), rel=2e-2)


def test_crossing_of_a_log_linear_waterfall():
    # FER = 10^-(x - 6): crosses 1e-3 at exactly x = 9.0
    rows = [{"es_n0_db": x, "fer": 10 ** -(x - 6.0)} for x in (8.7, 8.8, 8.9, 9.05, 9.2)]
    assert fs.interp_crossing(rows, "fer") == pytest.approx(9.0, abs=1e-4)
    assert fs.interp_crossing(rows[:2], "fer") is None  # never reaches 1e-3


def test_snr_scales_are_update_snr():
    nv, g = fs.snr_scales(9.2)
    assert nv == pytest.approx(10 ** -0.92, rel=1e-6) and g == pytest.approx(4 / nv, rel=1e-6)
    assert fs.snr_scales(40.0)[1] == 1000.0
