"""The benchmark's measurement support on the GPU: the VALU issue-rate
reference launch reports a rate a SIMD can issue (at most one wave
instruction per cycle) and scales with its iteration count."""
import pytest

from jxr_amd import HipDecoderSession

pytestmark = pytest.mark.gpu


def test_valu_calibration_reports_a_plausible_rate():
    s = HipDecoderSession(0)
    try:
        short, long_ = s.calibrate_valu(2000), s.calibrate_valu(20000)
        for r in (short, long_):
            assert r["simds"] > 0 and r["clock_khz"] > 0 and r["ms"] > 0
            rate = r["wave_instructions"] / (r["ms"] / 1e3) / (r["simds"] * r["clock_khz"] * 1e3)
            assert 0.05 < rate <= 1.0, rate
        assert long_["wave_instructions"] == 10 * short["wave_instructions"]
        assert long_["ms"] > 3 * short["ms"]
    finally:
        s.close()
