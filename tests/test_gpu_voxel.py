"""GPU voxelizer (pfdtd_voxelize) bit-exact against the oracle's voxelizer, its classification
for both boundary families, and the App's geometry path the reference's scripts drive
(python/testBench.py:42-149: JSON room -> initializeGeometryPy -> addSurfaceMaterials ->
runSimulation; App::initializeGeometryFromFile for VTK rooms)."""
import math
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle
from meshes import convex_shapes, l_room
from parallelfdtd_amd import geometry as G
from parallelfdtd_amd import native as N
from parallelfdtd_amd.solver import Solver

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "parallelfdtd_amd"))
LAM = math.sqrt(1.0 / 3.0)


def cases():
    out = []
    b = G.box_mesh((0, 0, 0), (1, 1, 1))
    out += [("box-0.1", b.vertices, b.indices, None, 0.1, (1, 1, 1)),
            ("box-0.3", b.vertices, b.indices, None, 0.3, (1, 1, 1))]
    b = G.box_mesh((0, 0, 0), (1.0, 0.8, 0.6))
    out.append(("box-faces", b.vertices, b.indices,
                G.layer_materials(b, {n: k for k, n in enumerate(b.layers)}), 0.07, (1, 1, 1)))
    room, table = G.box_mesh((0, 0, 0), (2.0, 1.6, 1.2)), G.box_mesh((0.5, 0.4, 0), (1.1, 1.2, 0.7))
    m = G.merge(room, table)
    out.append(("nested", m.vertices, m.indices, np.array([0] * 12 + [4] * 12, np.uint8), 0.1,
                (32, 4, 1)))
    for name, v, f in convex_shapes():
        out.append((name, np.asarray(v, np.float32), f, None, 0.037, (1, 1, 1)))
    v, f = l_room(0.9)
    out.append(("l-room", v, f, None, 0.1, (1, 1, 1)))
    bj = G.load_json(GOLDEN / "box.json")
    out.append(("box.json", bj.vertices, bj.indices,
                G.layer_materials(bj, {"floor": 0, "ceiling": 1, "walls": 2}), 0.1, (1, 1, 1)))
    h = G.load_json(GOLDEN / "larun_hytti.json")
    hm = G.layer_materials(h, {n: k for k, n in enumerate(h.layers)})
    for dx in (0.1, 0.05, 0.02):
        out.append((f"hytti-{dx}", h.vertices, h.indices, hm, dx, (32, 4, 1)))
    return out


@pytest.mark.parametrize("case", cases(), ids=lambda c: c[0])
def test_labels_and_domain_bit_exact(gpu, case):
    name, v, f, tmat, dx, pad = case
    ref = oracle.voxelize(v, f, dx, tmat, pad)
    with Solver() as s:
        lab = s.voxelize(v, f, dx, tmat, pad, want_labels=True)
        pos, mat = s.domain()
    assert lab.shape == ref.shape
    assert np.array_equal(lab, ref), f"{name}: {np.sum(lab != ref)} labels differ"
    pos_o, mat_o, _ = oracle.classify(ref)
    assert np.array_equal(pos, pos_o) and np.array_equal(mat, mat_o)


def test_centred_classification_of_voxelized_room(gpu):
    h = G.load_json(GOLDEN / "larun_hytti.json")
    ref = oracle.voxelize(h.vertices, h.indices, 0.05)
    with Solver() as s:
        s.set_update_type(N.SRL)
        s.voxelize(h.vertices, h.indices, 0.05)
        pos, mat = s.domain()
    pos_o, mat_o, _ = oracle.classify_kow(ref)
    assert np.array_equal(pos, pos_o) and np.array_equal(mat, mat_o)


def test_large_room(gpu):
    """The reference room at 1 cm (~6e7 nodes): labels bit-exact, one voxelization call."""
    h = G.load_json(GOLDEN / "larun_hytti.json")
    hm = G.layer_materials(h, {n: k for k, n in enumerate(h.layers)})
    ref = oracle.voxelize(h.vertices, h.indices, 0.01, hm, (32, 4, 1))
    with Solver() as s:
        lab = s.voxelize(h.vertices, h.indices, 0.01, hm, (32, 4, 1), want_labels=True)
    assert lab.size > 5e7
    assert np.array_equal(lab, ref)


def test_voxelize_rejects_bad_meshes(gpu):
    b = G.box_mesh((0, 0, 0), (1, 1, 1))
    bad = b.indices.copy()
    bad[0, 0] = 8
    with Solver() as s:
        for args in ((b.vertices, bad, 0.1), (b.vertices, b.indices, -1.0),
                     (b.vertices * 1e4, b.indices, 0.1)):
            with pytest.raises(N.PfdtdError) as ei:
                s.voxelize(*args)
            assert ei.value.code == N.E_ARG
        with pytest.raises(N.PfdtdError):
            s.voxelize(b.vertices, b.indices, 0.1, np.full(12, 255, np.uint8))


# ------------------------------------------------------------------------------------------
# App geometry path (libPyFDTD)
# ------------------------------------------------------------------------------------------
def _unique_order(values):
    """MaterialHandler::checkIfUnique: unique index = order of first occurrence."""
    seen, idx = [], []
    for v in values:
        if v not in seen:
            seen.append(v)
        idx.append(seen.index(v))
    return seen, np.array(idx, np.uint8)


def _oracle_app_responses(app, mesh, betas_per_tri, n_steps):
    uniq, tmat = _unique_order(list(betas_per_tri))
    dx = app.getDx()
    lab = oracle.voxelize(mesh.vertices, mesh.indices, dx, tmat, (32, 4, 1))
    assert (app.getNumElems(),) == (lab.size,)
    pos, mat, _ = oracle.classify(lab)
    ns, nr = app.getNumSources(), app.getNumReceivers()
    src = np.array([app.getSourceElementCoordinates(i) for i in range(ns)], np.int32)
    rec = np.array([app.getReceiverElementCoordinates(i) for i in range(nr)], np.int32)
    sig = np.array([[app.getSourceSample(i, n) for n in range(n_steps)] for i in range(ns)])
    table = np.repeat(np.array(uniq, np.float32)[:, None], 20, axis=1)
    return oracle.run(pos, mat, table, src, np.ones(ns, np.int32), sig, rec, n_steps)


def _geometry_app(mesh, betas, fs, n_steps, from_file=None):
    pf = pytest.importorskip("libPyFDTD")
    app = pf.App()
    app.initializeDevices()
    if from_file:
        app.initializeGeometryFromFile(str(from_file))
    else:
        app.initializeGeometryPy(mesh.indices.ravel().tolist(), mesh.vertices.ravel().tolist())
    app.setUpdateType(0)
    app.setNumSteps(n_steps)
    app.setSpatialFs(fs)
    app.setDouble(False)
    mats = np.repeat(np.asarray(betas, np.float32)[:, None], 20, axis=1)
    app.addSurfaceMaterials(mats.ravel().tolist(), len(betas), 20)
    app.addSource(1.2, 2.6, 1.3, 1, 1, 0)  # soft Gaussian
    app.addReceiver(2.3, 3.1, 1.6)
    app.addReceiver(0.6, 1.1, 0.7)
    return app


def test_app_json_room_matches_oracle(gpu):
    """testBench.py's flow on the reference's room (5 layers, distinct betas) at fs = 10 kHz."""
    mesh = G.load_json(GOLDEN / "larun_hytti.json")
    layer_beta = {"Layer0": 0.05, "window": 0.02, "floor": 0.3, "ceiling": 0.12, "table": 0.45}
    per_tri = np.zeros(mesh.n_triangles, np.float32)
    for name, b in layer_beta.items():
        per_tri[mesh.layers[name]] = b
    n_steps = 300
    app = _geometry_app(mesh, per_tri, 10000, n_steps)
    app.runSimulation()
    ref = _oracle_app_responses(app, mesh, per_tri, n_steps)
    got = np.array([app.getResponse(r) for r in range(2)])
    assert np.abs(ref).max() > 1e-4
    assert np.array_equal(got, ref)
    app.close()


def test_app_vtk_room_matches_oracle(gpu, tmp_path):
    """initializeGeometryFromFile (inches, rounded to 1 cm) equals the Python reader's mesh."""
    mesh = G.load_json(GOLDEN / "larun_hytti.json")
    G.write_vtk(tmp_path / "room.vtk", mesh)
    vtk_mesh = G.read_vtk(tmp_path / "room.vtk")
    per_tri = np.full(mesh.n_triangles, 0.1, np.float32)
    n_steps = 120
    app = _geometry_app(mesh, per_tri, 8000, n_steps, from_file=tmp_path / "room.vtk")
    app.runSimulation()
    ref = _oracle_app_responses(app, vtk_mesh, per_tri, n_steps)
    got = np.array([app.getResponse(r) for r in range(2)])
    assert np.array_equal(got, ref)
    app.close()
