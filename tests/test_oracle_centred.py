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


# --------------------------------------------------------------------------------------
# energy balance of the centred lossy boundary
# --------------------------------------------------------------------------------------
# At a centred boundary node the missing neighbour on each of its D = dir_x + dir_y + dir_z
# axes takes the value of the present one (S_boundary, kernels3d.cu:645-650), i.e. the node
# sits on a mirror plane. With weights w_i = 2^-D_i and edge weights W_ij = w_i c_ij (c_ij = 2
# when the edge runs along one of i's mirrored axes, else 1) the scheme's Laplacian
# (L p)_i = sum_j c_ij (p_j - p_i) is symmetric in a box room, and with lambda^2 = 1/3 the
# update (1 + beta) p^{n+1} = lambda^2 S + (beta - 1) p^{n-1} (:652) reads
#     p^{n+1} - 2 p^n + p^{n-1} = lambda^2 L p^n - beta (p^{n+1} - p^{n-1}).
# Multiplying by w_i (p^{n+1}_i - p^{n-1}_i) and summing gives, exactly,
#     E^{n+1/2} - E^{n-1/2} = -sum_i w_i beta_i (p^{n+1}_i - p^{n-1}_i)^2,
#     E^{n+1/2} = sum_i w_i (p^{n+1}_i - p^n_i)^2 + lambda^2 sum_{edges} W_ij dp^{n+1} dp^n,
# with beta_i = table[mat_i] * lambda * D_i (:639-640). A beta off by any factor, or taken per
# node instead of per mirrored axis, breaks the identity.
LAM = math.sqrt(1.0 / 3.0)


def _centred_room():
    """Box room whose six walls carry four materials (x walls 0 / 1, y walls 2, z walls 3)."""
    X, Y, Z = 12, 10, 9
    lab = rooms.box_room(X, Y, Z)
    lab[:, :, 0] = 1
    lab[:, :, X - 1] = 2
    lab[:, 0, :] = lab[:, Y - 1, :] = 3
    lab[0, :, :] = lab[Z - 1, :, :] = 4
    pos, mat, counts = oracle.classify_kow(lab)
    assert counts[2] == 0
    return pos, mat


def _weights(pos):
    inside = (pos >> 7).astype(bool)
    dirs = np.stack([(pos >> a) & 1 for a in range(3)]).astype(np.int64)  # [axis][z][y][x]
    D = dirs.sum(axis=0)
    w = np.where(inside, 2.0 ** (-D), 0.0)
    edges = []  # (axis, W for the edge between node and its + neighbour along axis)
    for ax in range(3):
        npax = 2 - ax  # numpy axis of x (0) / y (1) / z (2)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[npax] = slice(None, -1)
        hi[npax] = slice(1, None)
        both = inside[tuple(lo)] & inside[tuple(hi)]
        W_lo = w[tuple(lo)] * (1 + dirs[ax][tuple(lo)])
        W_hi = w[tuple(hi)] * (1 + dirs[ax][tuple(hi)])
        assert np.array_equal(W_lo[both], W_hi[both])  # symmetric edge weights
        edges.append((tuple(lo), tuple(hi), np.where(both, W_lo, 0.0)))
    return inside, D, w, edges


def _energy(p1, p0, w, edges):
    e = np.sum(w * (p1 - p0) ** 2)
    for lo, hi, W in edges:
        e += LAM * LAM * np.sum(W * (p1[lo] - p1[hi]) * (p0[lo] - p0[hi]))
    return e


def _fields(pos, mat, table, n, src, sig):
    _, P, Pp = oracle.run(pos, mat, table, src, np.ones(1, np.int32), sig, src[:0], n,
                          precision="f64", literal=oracle.SCHEME_KOWALCZYK, want_fields=True)
    return P, Pp


def test_centred_energy_conserved_rigid_walls():
    pos, mat = _centred_room()
    _, _, w, edges = _weights(pos)
    src = np.array([[4, 5, 3]], np.int32)
    sig = np.zeros((1, 60))
    sig[0, 0] = 1.0
    table = np.zeros((5, 20), np.float32)
    Es = []
    for n in range(2, 50, 8):
        P, Pp = _fields(pos, mat, table, n, src, sig)
        Es.append(_energy(P, Pp, w, edges))
    Es = np.array(Es)
    assert Es[0] > 0.1
    assert np.max(np.abs(Es - Es[0])) < 1e-12 * Es[0]


def test_centred_energy_balance_lossy_walls():
    pos, mat = _centred_room()
    inside, D, w, edges = _weights(pos)
    betas = np.array([0.08, 0.35, 0.6, 0.2, 0.45], np.float32)
    table = np.zeros((5, 20), np.float32)
    table[:, 0] = betas
    beta = betas.astype(np.float64)[mat] * LAM * D * inside  # beta = table * lambda * D
    assert set(np.unique(D[inside])) == {0, 1, 2, 3}
    src = np.array([[3, 4, 2]], np.int32)  # near a corner: faces, edges and corners absorb
    sig = np.zeros((1, 60))
    sig[0, 0] = 1.0
    prev = None
    for n in range(3, 40, 6):
        P0, Pm = _fields(pos, mat, table, n, src, sig)  # p^n, p^{n-1}
        P1, _ = _fields(pos, mat, table, n + 1, src, sig)
        e_hi, e_lo = _energy(P1, P0, w, edges), _energy(P0, Pm, w, edges)
        d2 = (P1 - Pm) ** 2
        loss = -np.sum(w * beta * d2)
        assert loss < 0
        assert abs((e_hi - e_lo) - loss) < 1e-11 * e_lo
        # the identity fails for a beta without the per-axis count D, or halved
        for wrong in (betas.astype(np.float64)[mat] * LAM * inside, 0.5 * beta):
            bad = -np.sum(w * wrong * d2)
            assert abs((e_hi - e_lo) - bad) > 1e-6 * abs(loss)
        if prev is not None:
            assert e_hi < prev
        prev = e_hi
