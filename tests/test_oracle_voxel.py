"""Pins of the oracle's triangle-mesh voxelizer (orc_voxelize).

The reference hands voxelization to the external hakarlss/Voxelizer (not vendored, no
version pinned; call site src/kernels/voxelizationUtils.cu:58-131), so there is no reference
output to match. What pins the definition instead:

* closed forms for axis-aligned rooms (walls on node planes and between them): grid dims, the
  air set, and the material of every boundary solid node (per-face materials);
* an independent inside test -- half-space tests for convex solids (rotated boxes,
  tetrahedra, an icosphere) and the generalized winding number (Van Oosterom-Strackee solid
  angles) for the reference's own room mesh python/Data/larun_hytti.json;
* the enclosed volume of the reference's meshes (box.json exactly, larun_hytti.json in the
  limit dx -> 0);
* invariance under triangle order, vertex rotation, winding and exact translation.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

import oracle
from parallelfdtd_amd import geometry as G
from meshes import convex_shapes, l_room

GOLDEN = Path(__file__).resolve().parent / "golden"
SUB = 4096


def half_open_count(L: float, dx: float) -> int:
    """Nodes at 0, dx, 2dx, ... inside [0, L)."""
    r = L / dx
    return int(math.ceil(r - 1e-9)) if abs(r - round(r)) > 1e-6 else int(round(r))


def node_coords(dims, dx):
    X, Y, Z = dims
    x = (np.arange(X) - 1) * dx
    y = (np.arange(Y) - 1) * dx
    z = (np.arange(Z) - 1) * dx
    return np.meshgrid(z, y, x, indexing="ij")  # (Z, Y, X) arrays of z, y, x


def box_air(dims, dx, boxes):
    """Union of half-open boxes [lo, hi) minus holes, on node centres (frame at bbox min)."""
    zz, yy, xx = node_coords(dims, dx)
    air = np.zeros(zz.shape, bool)
    for (lo, hi), sign in boxes:
        inside = ((xx >= lo[0] - 1e-9) & (xx < hi[0] - 1e-9) & (yy >= lo[1] - 1e-9) &
                  (yy < hi[1] - 1e-9) & (zz >= lo[2] - 1e-9) & (zz < hi[2] - 1e-9))
        air = air | inside if sign > 0 else air & ~inside
    return air


@pytest.mark.parametrize("hi,dx", [((1, 1, 1), 0.1), ((1, 2, 0.5), 0.1), ((1, 1, 1), 0.3),
                                   ((0.75, 0.5, 1.25), 0.0625), ((1.3, 0.7, 0.9), 0.11)])
def test_box_room_closed_form(hi, dx):
    m = G.box_mesh((0, 0, 0), hi)
    lab = oracle.voxelize(m.vertices, m.indices, dx)
    Z, Y, X = lab.shape
    for L, n in zip(hi, (X, Y, Z)):
        assert n == int(math.floor(L / dx + 1e-9)) + 3
    air = lab == 0
    assert air.sum() == np.prod([half_open_count(L, dx) for L in hi])
    assert np.array_equal(air, box_air((X, Y, Z), dx, [(((0, 0, 0), hi), 1)]))


def test_box_padding_to_block_multiples():
    m = G.box_mesh((0, 0, 0), (1, 1, 1))
    lab = oracle.voxelize(m.vertices, m.indices, 0.1, pad=(32, 4, 1))
    assert lab.shape == (13, 16, 32)
    ref = oracle.voxelize(m.vertices, m.indices, 0.1)
    assert np.array_equal(lab[:, :13, :13], ref)
    assert np.all(lab[:, 13:, :] == 1) and np.all(lab[:, :, 13:] == 1)


FACE_OF_FIRST_AIR = ["+x", "-x", "+y", "-y", "+z", "-z"]  # first air neighbour -x means +x wall


def expected_face_labels(air, face_label):
    """Brute force: solid voxel -> label of the face across its first air neighbour."""
    Z, Y, X = air.shape
    lab = np.where(air, 0, 1).astype(np.uint8)
    pad = np.pad(air, 1, constant_values=False)
    shifts = [(0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)]
    done = air.copy()
    for (dz, dy, dx_), face in zip(shifts, FACE_OF_FIRST_AIR):
        nb = pad[1 + dz:1 + dz + Z, 1 + dy:1 + dy + Y, 1 + dx_:1 + dx_ + X]
        hit = nb & ~done
        lab[hit] = face_label[face]
        done |= hit
    return lab


@pytest.mark.parametrize("dx", [0.1, 0.07])
def test_box_face_materials(dx):
    m = G.box_mesh((0, 0, 0), (1.0, 0.8, 0.6))
    names = list(m.layers)
    tmat = G.layer_materials(m, {n: k for k, n in enumerate(names)})
    lab = oracle.voxelize(m.vertices, m.indices, dx, tmat)
    exp = expected_face_labels(lab == 0, {n: k + 1 for k, n in enumerate(names)})
    assert np.array_equal(lab, exp)
    assert set(np.unique(lab)) == {0, 1, 2, 3, 4, 5, 6}


def test_box_json_fixture():
    m = G.load_json(GOLDEN / "box.json")
    assert m.n_triangles == 12 and set(m.layers) == {"floor", "ceiling", "walls"}
    assert G.mesh_volume(m) == pytest.approx(1.0)
    tmat = G.layer_materials(m, {"floor": 0, "ceiling": 1, "walls": 2})
    lab = oracle.voxelize(m.vertices, m.indices, 0.1, tmat)
    assert lab.shape == (13, 13, 13) and (lab == 0).sum() == 1000
    air = lab == 0
    exp = expected_face_labels(air, {"-x": 3, "+x": 3, "-y": 3, "+y": 3, "-z": 1, "+z": 2})
    assert np.array_equal(lab, exp)


def test_nested_box_is_solid_and_keeps_its_material():
    room = G.box_mesh((0, 0, 0), (2.0, 1.6, 1.2))
    table = G.box_mesh((0.5, 0.4, 0.0), (1.1, 1.2, 0.7))  # standing on the floor
    m = G.merge(room, table)
    tmat = np.array([0] * 12 + [4] * 12, np.uint8)
    dx = 0.1
    lab = oracle.voxelize(m.vertices, m.indices, dx, tmat)
    Z, Y, X = lab.shape
    air = box_air((X, Y, Z), dx, [(((0, 0, 0), (2.0, 1.6, 1.2)), 1),
                                  (((0.5, 0.4, 0.0), (1.1, 1.2, 0.7)), -1)])
    assert np.array_equal(lab == 0, air)
    zz, yy, xx = node_coords((X, Y, Z), dx)
    in_table = (xx > 0.45) & (xx < 1.05) & (yy > 0.35) & (yy < 1.15) & (zz > -0.05) & (zz < 0.65)
    pad = np.pad(air, 1)
    touches_air = np.zeros_like(air)
    for ax in range(3):
        for d in (-1, 1):
            touches_air |= np.roll(pad, d, axis=ax)[1:-1, 1:-1, 1:-1]
    assert np.all(lab[in_table & touches_air & ~air] == 5)
    assert np.all(lab[~in_table & touches_air & ~air] == 1)


@pytest.mark.parametrize("name,v,f", convex_shapes(), ids=lambda x: x if isinstance(x, str) else "")
def test_convex_solids_match_half_space_test(name, v, f):
    dx = 0.037
    v = np.asarray(v, np.float32)
    lab = oracle.voxelize(v, f, dx)
    Z, Y, X = lab.shape
    vr = (v - v.min(0)).astype(np.float64)
    zz, yy, xx = node_coords((X, Y, Z), dx)
    P = np.stack([xx, yy, zz], -1).reshape(-1, 3)
    a, b, c = vr[f[:, 0]], vr[f[:, 1]], vr[f[:, 2]]
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    d = (P @ n.T) - np.einsum("ij,ij->i", n, a)[None, :]  # signed distances, outward positive
    inside = np.all(d < 0, axis=1)
    clear = np.min(np.abs(d), axis=1) > 1e-3 * dx
    got = (lab == 0).reshape(-1)
    assert inside.sum() > 100
    assert np.array_equal(got[clear], inside[clear]), f"{name}: {np.sum(got[clear] != inside[clear])}"
    # nodes on or next to a face plane (outside the clear set) stay within one voxel of it
    assert np.all(np.min(np.abs(d[~clear]), axis=1) < dx)


def winding_number(P, v, f, chunk=4096):
    w = np.zeros(len(P))
    for s in range(0, len(P), chunk):
        p = P[s:s + chunk, None, :]
        a, b, c = v[f[:, 0]][None] - p, v[f[:, 1]][None] - p, v[f[:, 2]][None] - p
        la, lb, lc = (np.linalg.norm(x, axis=2) for x in (a, b, c))
        det = np.einsum("nij,nij->ni", a, np.cross(b, c))
        den = (la * lb * lc + np.einsum("nij,nij->ni", a, b) * lc +
               np.einsum("nij,nij->ni", a, c) * lb + np.einsum("nij,nij->ni", b, c) * la)
        w[s:s + chunk] = (2 * np.arctan2(det, den)).sum(1) / (4 * np.pi)
    return w


@pytest.mark.parametrize("dx", [0.2, 0.1])
def test_reference_room_matches_winding_number(dx):
    m = G.load_json(GOLDEN / "larun_hytti.json")
    lab = oracle.voxelize(m.vertices, m.indices, dx)
    Z, Y, X = lab.shape
    # the voxelizer sees fixed-point vertices: evaluate the winding number on exactly those
    vr = m.vertices - m.vertices.min(0)
    q = np.rint(vr.astype(np.float64) * SUB / np.float64(np.float32(dx)))
    vq = q / SUB * np.float64(np.float32(dx))
    zz, yy, xx = node_coords((X, Y, Z), np.float64(np.float32(dx)))
    P = np.stack([xx, yy, zz], -1).reshape(-1, 3)
    # nodes on the surface have no defined inside: drop those whose winding number differs a
    # hair to either side (along a direction no wall contains)
    e = 1e-5 * dx * np.array([1.0, 2 ** 0.5, 3 ** 0.5]) / 6 ** 0.5
    f = m.indices.astype(np.int64)
    w, wp, wm = (winding_number(P + d, vq, f) for d in (0 * e, e, -e))
    sure = (np.abs(w - np.rint(w)) < 1e-6) & (np.rint(wp) == np.rint(w)) & (np.rint(wm) == np.rint(w))
    assert sure.mean() > 0.9  # floor and two walls lie on node planes
    assert set(np.unique(np.rint(w[sure]))) <= {0.0, 1.0}
    got = (lab == 0).reshape(-1)
    assert np.array_equal(got[sure], np.rint(w[sure]) == 1)


def test_reference_room_volume_converges():
    m = G.load_json(GOLDEN / "larun_hytti.json")
    V = G.mesh_volume(m)
    assert V == pytest.approx(48.2635, rel=1e-4)
    errs = []
    for dx in (0.1, 0.05, 0.025):
        lab = oracle.voxelize(m.vertices, m.indices, dx)
        errs.append(abs((lab == 0).sum() * dx ** 3 - V) / V)
    assert errs[0] < 0.06 and errs[1] < 0.02 and errs[2] < 0.01
    assert errs[2] < errs[0]


def test_l_shaped_room_closed_form():
    v, f = l_room(0.9)
    m = G.Mesh(v, f)
    assert G.mesh_volume(m) == pytest.approx(3 * 0.9)
    dx = 0.1
    lab = oracle.voxelize(v, f, dx)
    Z, Y, X = lab.shape
    air = box_air((X, Y, Z), dx, [(((0, 0, 0), (2, 1, 0.9)), 1), (((0, 0, 0), (1, 2, 0.9)), 1)])
    assert np.array_equal(lab == 0, air)


def test_invariances():
    room = G.box_mesh((0, 0, 0), (1.5, 1.25, 1.0))
    table = G.box_mesh((0.25, 0.5, 0.0), (0.75, 1.0, 0.5))
    m = G.merge(room, table)
    tmat = np.array([k // 2 for k in range(12)] + [6] * 12, np.uint8)
    dx = 0.0625
    ref = oracle.voxelize(m.vertices, m.indices, dx, tmat)
    # vertex rotation inside each triangle and reversed winding
    for perm in ([1, 2, 0], [2, 0, 1], [0, 2, 1]):
        assert np.array_equal(oracle.voxelize(m.vertices, m.indices[:, perm], dx, tmat), ref)
    # exact translation (binary fractions)
    assert np.array_equal(oracle.voxelize(m.vertices + np.float32(3.25), m.indices, dx, tmat), ref)
    # triangle order, when each link is crossed by one triangle
    rng = np.random.default_rng(1)
    order = rng.permutation(m.n_triangles)
    assert np.array_equal(oracle.voxelize(m.vertices, m.indices[order], dx, tmat[order]), ref)


def test_rejects_bad_input():
    m = G.box_mesh((0, 0, 0), (1, 1, 1))
    with pytest.raises(ValueError):
        oracle.voxelize(m.vertices, m.indices, 0.0)
    bad = m.indices.copy()
    bad[3, 1] = 99
    with pytest.raises(ValueError):
        oracle.voxelize(m.vertices, bad, 0.1)
    with pytest.raises(ValueError):
        oracle.voxelize(m.vertices, m.indices, 0.1, np.full(12, 255, np.uint8))
    with pytest.raises(ValueError):
        oracle.voxelize(m.vertices * 1e4, m.indices, 0.1)  # grid above 32768 per axis


def test_json_and_vtk_readers(tmp_path):
    m = G.load_json(GOLDEN / "larun_hytti.json")
    assert m.n_triangles == 166 and len(m.vertices) == 87
    assert {k: len(v) for k, v in m.layers.items()} == {"Layer0": 98, "window": 4, "floor": 26,
                                                        "ceiling": 26, "table": 12}
    G.write_vtk(tmp_path / "room.vtk", m)
    r = G.read_vtk(tmp_path / "room.vtk")
    assert np.array_equal(r.indices, m.indices)
    # inches -> metres, rounded to whole centimetres
    assert np.allclose(r.vertices, np.round(m.vertices.astype(np.float64), 2), atol=1e-6)
    assert np.allclose(r.vertices * 100, np.rint(r.vertices * 100), atol=1e-4)
