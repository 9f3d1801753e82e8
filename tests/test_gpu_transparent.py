"""Transparent-source signals on the device (pfdtd_transparent_signals) against the oracle's
restatement of getTransparentSourceSample(+Double) (SimulationParameters.cpp:300-310,
:350-360): bit-exact in FP32 and FP64, and through FDTD::App."""
import ctypes as C

import numpy as np
import pytest

import oracle
from parallelfdtd_amd.solver import Solver

pytestmark = pytest.mark.gpu


def oracle_transparent(data, h, prec):
    L = oracle.lib()
    n = data.shape[1]
    out = np.empty_like(data)
    for s in range(data.shape[0]):
        row = np.ascontiguousarray(data[s])
        for k in range(n):
            if prec == "f32":
                out[s, k] = L.orc_transparent_sample_f(3, k, 1000, row.ctypes.data, n, h.ctypes.data, h.size)
            else:
                out[s, k] = L.orc_transparent_sample_d(3, k, 1000, row.ctypes.data, n, h.ctypes.data, h.size)
    return out


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_transparent_signals_bit_exact(gpu, prec):
    rng = np.random.default_rng(17)
    n = 700
    data = rng.uniform(-1, 1, (3, n)).astype(np.float32 if prec == "f32" else np.float64)
    h = np.zeros(400, np.float32)
    h[2::2] = (-1.0 / 3.0) * (0.97 ** np.arange(199))  # a decaying grid response, h[2] = -1/3
    with Solver() as s:
        got = s.transparent_signals(data, h)
    ref = oracle_transparent(data, h, prec)
    assert np.array_equal(got, ref)
    # s_T(3) = 1/3 for an impulse at step 1 (SimulationParametersTest.cpp:196-207)
    imp = np.zeros((1, 8), np.float32)
    imp[0, 1] = 1.0
    with Solver() as s:
        t = s.transparent_signals(imp, h)
    assert t[0, 3] == np.float32(1.0 / 3.0)
