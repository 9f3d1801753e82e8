"""The CPU oracle against the committed fixtures of tests/golden/ (an independent numpy
evaluation written by tests/golden/make_golden.py): bit-exact classification, responses and
final fields in FP64 and FP32, including the literal quirk mode, and with FP32 arithmetic over
binary16 storage in the plain, difference and sum-difference forms."""
from pathlib import Path

import numpy as np
import pytest

import oracle

GOLDEN = sorted((Path(__file__).resolve().parent / "golden").glob("*.npz"))


@pytest.mark.parametrize("path", GOLDEN, ids=[p.stem for p in GOLDEN])
def test_oracle_matches_golden(path):
    d = np.load(path, allow_pickle=False)
    pos, mat, _ = oracle.classify(d["labels"])
    assert np.array_equal(pos, d["pos"]) and np.array_equal(mat, d["mat"])
    for prec, key in (("f64", "64"), ("f32", "32")):
        sig = d["signals"] if prec == "f64" else d["signals"].astype(np.float32)
        r, P, Pp = oracle.run(pos, mat, d["table"], d["src"], d["kinds"], sig, d["rec"],
                              int(d["n_steps"]), precision=prec, octave=int(d["octave"]),
                              literal=int(d["literal"]), lam=float(d["lam"]), want_fields=True)
        assert np.array_equal(r, d["resp" + key])
        assert np.array_equal(P, d["P" + key]) and np.array_equal(Pp, d["Pp" + key])
    sig = d["signals"].astype(np.float32)
    for storage, key, fields in ((oracle.STORE_F16, "s16", ("P_s16", "Pp_s16")),
                                 (oracle.STORE_F16_DIFF, "d16", ("P_d16", "D_d16")),
                                 (oracle.STORE_F16_SUMDIFF, "sd16", ("S_sd16", "D_sd16"))):
        r, A, B = oracle.run(pos, mat, d["table"], d["src"], d["kinds"], sig, d["rec"],
                             int(d["n_steps"]), octave=int(d["octave"]), literal=int(d["literal"]),
                             lam=float(d["lam"]), storage=storage, scale_exp=int(d["scale16"]),
                             want_fields=True)
        assert np.abs(r).max() > 0.05
        assert np.array_equal(r, d["resp_" + key]), key
        assert np.array_equal(A, d[fields[0]]) and np.array_equal(B, d[fields[1]]), key


def test_golden_fixtures_present():
    assert len(GOLDEN) >= 4
