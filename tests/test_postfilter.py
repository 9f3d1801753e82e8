"""Response post-filters (matlab/functions/FDTDpostFilter.m, python/FDTDfilter.py): pinned by
the DC blocker's own inverse relation in dcblock.m, its gain at the cut-on, and the low-pass's DC
gain, linear-phase delay and stop-band."""
import math

import numpy as np
import pytest
from scipy.signal import freqz

from parallelfdtd_amd.postfilter import dc_block_coefficient, fdtd_filter, post_filter


@pytest.mark.parametrize("fc,fs", [(5.0, 40000.0), (20.0, 8000.0), (100.0, 48000.0)])
def test_dc_block_coefficient_inverse_relation(fc, fs):
    a = dc_block_coefficient(fc, fs)
    # dcblock.m's inverse: Fc = atan(sqrt(3)(1-a^2)/(1+4a+a^2))/pi
    Fc = math.atan(math.sqrt(3) * (1 - a * a) / (1 + 4 * a + a * a)) / math.pi
    assert Fc == pytest.approx(2 * fc / fs, rel=1e-12)
    # the gain-normalised blocker (1+a)/2 (1 - z^-1)/(1 - a z^-1) is at -6 dB (1/2) at fc
    w, h = freqz([1.0, -1.0], [1.0, -a], worN=[math.pi * 2 * fc / fs])
    assert abs(h[0]) * (1 + a) / 2 == pytest.approx(0.5, rel=1e-9)


def test_post_filter_passband_delay_and_stopband():
    fs, n = 40000.0, 4000
    t = np.arange(n)
    low = np.sin(2 * np.pi * 0.02 * t)           # 0.02 fs: in the pass band (0.1 fs)
    nyq = (-1.0) ** t                            # Nyquist: stop band
    # the low-pass alone: linear phase (delay order/2 = 100 samples), unit pass-band gain
    y = post_filter(np.stack([low, nyq]), fs, frac=0.2, dc_cut_on_hz=None)
    np.testing.assert_allclose(y[0, 1000:3000], low[900:2900], atol=2e-3)
    assert np.abs(y[1, 1000:]).max() < 1e-3
    # with the 5 Hz DC blocker the 800 Hz tone keeps its amplitude to ~1 %
    y = post_filter(low[None, :], fs, frac=0.2)
    assert np.abs(y[0, 1000:3000]).max() == pytest.approx(1.0, abs=2e-2)


def test_post_filter_blocks_dc():
    fs = 40000.0
    step = np.ones(400000)
    y = post_filter(step[None, :], fs, frac=0.2)[0]
    assert abs(y[-1]) < 1e-3 * abs(y[150])


def test_fdtd_filter_dc_gain_and_cutoff():
    taps_in = np.zeros(1000)
    taps_in[0] = 1.0
    h = fdtd_filter(taps_in, 40000.0, 0.1)[:200]
    assert h.sum() == pytest.approx(1.0, rel=1e-9)   # firwin scales to unit DC gain
    w, H = freqz(h, worN=[2 * np.pi * 0.3])           # 0.3 fs, well past the 0.1 fs cutoff
    assert abs(H[0]) < 10 ** (-60 / 20)
