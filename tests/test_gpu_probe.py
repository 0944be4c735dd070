"""The measured ceilings behind the SpMV roofline (spira_probe_ceilings)."""
import pytest

import spira_amd


@pytest.mark.gpu
def test_probe_ceilings_are_plausible_rates():
    dev = spira_amd.Device(0)
    g, b = dev.probe_ceilings()
    print(f"L2 gathers {g / 1e9:.1f} G requests/s, streaming read {b / 1e12:.3f} TB/s")
    # MI355X: hundreds of G L2 requests/s; a streaming read below the 8 TB/s peak
    assert 20e9 < g < 5e12
    assert 0.5e12 < b < 8.0e12
    dev.close()
