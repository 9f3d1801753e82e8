"""Pins of the oracle's centred-boundary update (update type SRL, the Kowalczyk scheme:
fdtd3dStdKowalczykMaterials, src/kernels/kernels3d.cu:609-665; classification toKowalczyk,
src/kernels/cudaMesh.cu:363-480).

* The code -> position-byte table equals the reference's, transcribed below as data.
* Label classification gives, for every air-neighbour configuration a voxelizer code names,
  the byte of that code.
* Image principle: with beta = 0 a centred wall is a mirror, so a room equals the half of a
  twice-as-long room with mirrored sources (for each axis, x, y and z).
* Away from walls (inside the light cone of the run) the centred and forward schemes are the
  same leapfrog.
"""
import math

import numpy as np
import pytest

import oracle
from parallelfdtd_amd import rooms

# toKowalczyk byte per voxelizer code 0..27 (cudaMesh.cu:363-480 with SIGN_Z 0x40, SIGN_Y 0x20,
# SIGN_X 0x10, DIR_Z 0x04, DIR_Y 0x02, DIR_X 0x01, CENTERED_MASK 0x80, cudaMesh.h:35-44).
REF_BYTES = [0x00, 0xC7, 0xD7, 0xE7, 0xF7, 0x87, 0x97, 0xA7, 0xB7, 0xC6, 0xE6, 0xC5, 0xD5, 0x86,
             0xA6, 0x85, 0x95, 0x83, 0x93, 0xA3, 0xB3, 0xC4, 0xA2, 0x82, 0x91, 0x81, 0x84, 0x80]

# Air directions each code's comment lists (L/R = -x/+x, I/O = -y/+y, D/U = -z/+z).
REF_AIR = {1: "DLI", 2: "DRI", 3: "DLO", 4: "DRO", 5: "ULI", 6: "URI", 7: "ULO", 8: "URO",
           9: "DLRI", 10: "DLRO", 11: "DLIO", 12: "DRIO", 13: "ULRI", 14: "ULRO", 15: "ULIO",
           16: "URIO", 17: "UDLI", 18: "UDRI", 19: "UDLO", 20: "UDRO", 21: "LRIOD",
           22: "LRODU", 23: "LRIDU", 24: "RIODU", 25: "LIODU", 26: "LRIOU", 27: "LRIODU"}
OFFSET = {"L": (-1, 0, 0), "R": (1, 0, 0), "I": (0, -1, 0), "O": (0, 1, 0), "D": (0, 0, -1),
          "U": (0, 0, 1)}


def test_code_table_matches_reference():
    L = oracle.lib()
    got = [L.orc_to_kowalczyk(k, None) for k in range(28)]
    assert got == REF_BYTES
    assert L.orc_to_kowalczyk(31, None) == 31  # codes past 27 match no branch and stay


@pytest.mark.parametrize("code", range(1, 28))
def test_label_classification_per_code(code):
    """A 5x5x5 grid whose centre voxel has exactly the code's air neighbours."""
    lab = np.ones((5, 5, 5), np.uint8) * 3  # solid, material 2
    lab[2, 2, 2] = 0
    for d in REF_AIR[code]:
        dx, dy, dz = OFFSET[d]
        lab[2 + dz, 2 + dy, 2 + dx] = 0
    pos, mat, counts = oracle.classify_kow(lab)
    assert pos[2, 2, 2] == REF_BYTES[code]
    assert mat[2, 2, 2] == (0 if code == 27 else 2)


def test_degenerate_voxels_counted():
    lab = rooms.box_room(9, 7, 7)
    lab[3, 3, 1:8] = 1   # a wall with a one-voxel slit leaves voxels with no air on an axis
    lab[3, 3, 4] = 0
    _, _, counts = oracle.classify_kow(lab)
    assert counts[2] >= 1


def _mirror_case(axis: int, n: int = 9, m: int = 8, steps: int = 40, precision="f64"):
    """Room A: n inside cells along `axis`; room B: 2n-1 cells, A mirrored about its first
    inside plane; beta = 0 (rigid centred walls everywhere)."""
    dims_a = [m + 2, m + 3, m + 1]
    dims_a[axis] = n + 2
    dims_b = list(dims_a)
    dims_b[axis] = 2 * n + 1
    lab_a = rooms.box_room(*dims_a)
    lab_b = rooms.box_room(*dims_b)
    src_a = np.array([[3, 4, 3], [5, 2, 4]], np.int32)
    rec_a = np.array([[1, 1, 1], [2, 3, 2], [4, 5, 5], [3, 2, 6]], np.int32)
    rec_a[:, axis] = [1, 2, 5, n]
    src_a[:, axis] = [1, 4]

    def to_b(p, mirror=False):
        q = p.copy()
        q[:, axis] = (n + 1 - p[:, axis]) if mirror else (n - 1 + p[:, axis])
        return q

    src_b = np.concatenate([to_b(src_a), to_b(src_a[1:], mirror=True)])
    sig = np.stack([rooms.gaussian_pulse(steps, 8.0, 2.0), -0.5 * rooms.gaussian_pulse(steps, 12.0, 3.0)])
    sig_b = np.concatenate([sig, sig[1:]])
    table = np.zeros((1, 20), np.float32)
    out = []
    for lab, src, s, rec in ((lab_a, src_a, sig, rec_a), (lab_b, src_b, sig_b, to_b(rec_a))):
        pos, mat, counts = oracle.classify_kow(lab)
        assert counts[2] == 0
        kinds = np.ones(len(src), np.int32)
        out.append(oracle.run(pos, mat, table, src, kinds, s if precision == "f64" else s.astype(np.float32),
                              rec, steps, precision=precision, literal=oracle.SCHEME_KOWALCZYK))
    return out


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_centred_wall_is_a_mirror(axis):
    a, b = _mirror_case(axis)
    assert np.abs(a).max() > 1e-3
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-12 * np.abs(a).max())


def test_centred_wall_is_a_mirror_fp32():
    a, b = _mirror_case(2, precision="f32")
    np.testing.assert_allclose(a, b, rtol=0, atol=2e-6 * np.abs(a).max())


def test_air_region_matches_forward_scheme():
    """Source and receivers more than `steps` cells from every wall: both schemes agree."""
    steps = 12
    lab = rooms.box_room(31, 31, 31)
    src = np.array([[15, 15, 15]], np.int32)
    rec = np.array([[15, 15, 15], [17, 14, 15], [12, 15, 18]], np.int32)
    sig = rooms.gaussian_pulse(steps, 3.0, 1.0)[None, :]
    table = np.full((1, 20), 0.3, np.float32)
    pos_f, mat_f, _ = oracle.classify(lab)
    pos_k, mat_k, _ = oracle.classify_kow(lab)
    for prec, tol in (("f64", 1e-13), ("f32", 1e-6)):
        s = sig if prec == "f64" else sig.astype(np.float32)
        f = oracle.run(pos_f, mat_f, table, src, np.ones(1, np.int32), s, rec, steps, precision=prec)
        k = oracle.run(pos_k, mat_k, table, src, np.ones(1, np.int32), s, rec, steps, precision=prec,
                       literal=oracle.SCHEME_KOWALCZYK)
        assert np.abs(f).max() > 0.05
        np.testing.assert_allclose(k, f, rtol=0, atol=tol * np.abs(f).max())


def test_absorbing_walls_lose_energy_relative_to_rigid():
    """beta > 0 on the walls: the late response is smaller than with rigid walls."""
    steps = 300
    lab = rooms.box_room(14, 12, 10)
    src = np.array([[4, 5, 4]], np.int32)
    rec = np.array([[9, 6, 6]], np.int32)
    sig = rooms.gaussian_pulse(steps, 10.0, 3.0)[None, :]
    pos, mat, _ = oracle.classify_kow(lab)
    e = []
    for beta in (0.0, 0.5):
        table = np.full((1, 20), beta, np.float32)
        r = oracle.run(pos, mat, table, src, np.ones(1, np.int32), sig, rec, steps, precision="f64",
                       literal=oracle.SCHEME_KOWALCZYK)
        late = r[0, steps // 2:]
        e.append(float(np.sum((late - late.mean()) ** 2)))
    assert e[1] < 0.5 * e[0]
