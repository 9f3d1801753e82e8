"""Response post-filters of the reference's tool chain, for impulse responses returned by the
solver (host-side signal conditioning; the responses are short, [n_rec][n_steps]).

* post_filter -- matlab/functions/FDTDpostFilter.m:1-31: an order-200 linear-phase FIR
  low-pass (window design with a Hamming window, cutoff `frac` of the Nyquist rate, unit DC
  gain, as MATLAB's fir1) followed by a one-pole DC blocker with a 5 Hz cut-on.
  matlab/testBench.m:204 uses frac = 0.2: the leapfrog's usable band.
* fdtd_filter -- python/FDTDfilter.py:13-19: a 200-tap FIR low-pass, Dolph-Chebyshev window
  with 70 dB ripple, cutoff `normalized_cutoff * sfs` Hz.
* dc_block_coefficient -- the pole a of y[n] = x[n] - x[n-1] + a y[n-1] whose -3 dB cut-on is
  fc: a = (sqrt(3) - 2 sin(pi Fc)) / (sin(pi Fc) + sqrt(3) cos(pi Fc)), Fc = 2 fc / fs (the
  standard DC-blocker design used by matlab/functions/FileExchange/dcblock.m).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.signal import firwin, lfilter


def dc_block_coefficient(fc: float, fs: float) -> float:
    Fc = 2.0 * fc / fs
    if not 0.0 <= Fc <= 1.0 / 3.0:
        raise ValueError("cut-on frequency must lie in [0, fs/6]")
    s, c = math.sin(math.pi * Fc), math.cos(math.pi * Fc)
    return (math.sqrt(3.0) - 2.0 * s) / (s + math.sqrt(3.0) * c)


def post_filter(ir: np.ndarray, fs: float, frac: float = 0.2, order: int = 200,
                dc_cut_on_hz: float | None = 5.0) -> np.ndarray:
    """Low-pass (FIR, order `order`, cutoff `frac` x Nyquist) then DC block (skipped when
    dc_cut_on_hz is None), along the last axis (one response per row)."""
    x = np.asarray(ir, np.float64)
    b_low = firwin(order + 1, frac, window="hamming")
    y = lfilter(b_low, 1.0, x, axis=-1)
    if dc_cut_on_hz is None:
        return y
    a = dc_block_coefficient(dc_cut_on_hz, fs)
    return lfilter([1.0, -1.0], [1.0, -a], y, axis=-1)


def fdtd_filter(x: np.ndarray, sfs: float, normalized_cutoff: float, numtaps: int = 200,
                ripple_db: float = 70.0, axis: int = 0) -> np.ndarray:
    """FIR low-pass of python/FDTDfilter.py (Chebyshev window), along `axis`."""
    nyq = sfs / 2.0
    taps = firwin(numtaps, sfs * normalized_cutoff / nyq, window=("chebwin", ripple_db))
    return lfilter(taps, 1.0, np.asarray(x, np.float64), axis=axis)
