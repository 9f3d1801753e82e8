"""The FDTD::App facade through the libPyFDTD module (the reference's Python surface,
src/AppPy.cpp:96-134): host semantics on the CPU against the oracle, and reference-style
runs on the GPU against the oracle."""
import math
import os
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle
from parallelfdtd_amd import rooms

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "parallelfdtd_amd"))
pf = pytest.importorskip("libPyFDTD")

LAM = math.sqrt(1.0 / 3.0)


def test_python_surface_has_reference_methods():
    # the 25 names bound by AppPy.cpp:107-132
    names = ["initializeDevices", "initializeGeometryFromFile", "initializeGeometryPy",
             "setLayerIndices", "addSource", "addSourceDataFloat", "addSourceDataDouble",
             "addReceiver", "addSurfaceMaterials", "setSpatialFs", "setNumSteps", "setUpdateType",
             "setUniform", "runVisualization", "runSimulation", "runCapture", "setUniformMaterial",
             "getResponse", "getResponseDouble", "forcePartitionTo", "addSliceToCapture",
             "setDouble", "setCapturedB", "close", "getMvox", "getNumElems"]
    app = pf.App()
    assert all(hasattr(app, n) for n in names)


def test_parameter_vector_and_dx():
    app = pf.App()
    app.setUpdateType(0)  # lambda = sqrt(1/3), as SimulationParameters::setUpdateType sets it
    f = np.zeros(4, np.float32)
    oracle.lib().orc_param_vector_f(LAM, 0, f.ctypes.data)
    assert np.array_equal(np.array(app.getParameterVector(), np.float32), f)
    app.setSpatialFs(2000)
    assert app.getDx() == oracle.lib().orc_dx(344.0, 2000, LAM)


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("signal", [0, 1, 2, 3])
def test_source_samples_match_oracle(kind, signal):
    app = pf.App()
    app.setUpdateType(0)
    app.setSpatialFs(48000)
    data = np.linspace(-1, 1, 37).astype(np.float32)
    app.addSourceDataFloat(data.tolist(), len(data), 1)
    app.addSourceDataDouble(data.astype(np.float64).tolist(), len(data), 1)
    h = np.array([0.0, 0.0, -1 / 3, 0.0, 1 / 9, 0.0, -0.0617], np.float32)
    app.setGridIr(h.tolist())
    app.addSource(0.5, 0.25, 0.75, kind, signal, 0)
    L = oracle.lib()
    for n in range(60):
        if kind == 2:
            ef = L.orc_transparent_sample_f(signal, n, 48000, data.ctypes.data, len(data), h.ctypes.data, len(h))
            d64 = data.astype(np.float64)
            ed = L.orc_transparent_sample_d(signal, n, 48000, d64.ctypes.data, len(data), h.ctypes.data, len(h))
        else:
            ef = L.orc_source_sample_f(signal, n, 48000, data.ctypes.data, len(data))
            d64 = data.astype(np.float64)
            ed = L.orc_source_sample_d(signal, n, 48000, d64.ctypes.data, len(data))
        assert np.float32(app.getSourceSample(0, n)) == np.float32(ef)
        assert app.getSourceSampleDouble(0, n) == ed


def test_element_coordinates_match_oracle():
    app = pf.App()
    app.setUpdateType(0)
    app.setSpatialFs(20000)
    pts = [(0.31, 0.47, 0.123), (1.0, 0.0, 2.5)]
    for p in pts:
        app.addSource(*p, 0, 0, 0)
        app.addReceiver(*p)
    out = np.zeros(3, np.int32)
    for i, p in enumerate(pts):
        oracle.lib().orc_element_idx(*p, 20000, 344.0, LAM, 1, out.ctypes.data)
        assert app.getSourceElementCoordinates(i) == tuple(out)
        assert app.getReceiverElementCoordinates(i) == tuple(out)


def test_out_of_range_and_missing_domain():
    app = pf.App()
    with pytest.raises(IndexError):
        app.getSourceSample(3, 0)
    with pytest.raises(RuntimeError):
        app.runVisualization()


# ------------------------------------------------------------------------------------------
# GPU: reference-style runs
# ------------------------------------------------------------------------------------------
def _room_app(fs, n_steps, double=False, box=(40, 32, 24)):
    lab = rooms.box_room(*box, floor=0, ceiling=1, walls=1)
    lab[8:14, 10:20, 12:22] = 3  # an obstacle of material 2
    betas = [0.02, 0.3, 0.55]
    app = pf.App()
    app.initializeDevices()
    app.setUpdateType(0)
    app.setSpatialFs(fs)
    app.setNumSteps(n_steps)
    app.setDouble(double)
    for b in betas:
        app.addSurfaceMaterials([b] * 20, 1, 20)
    app.initializeVoxelLabels(lab)
    return app, lab, np.array(betas, np.float32)


def _oracle_responses(app, lab, betas, n_steps, precision):
    pos, mat, _ = oracle.classify(lab)
    ns, nr = app.getNumSources(), app.getNumReceivers()
    src = np.array([app.getSourceElementCoordinates(i) for i in range(ns)], np.int32)
    rec = np.array([app.getReceiverElementCoordinates(i) for i in range(nr)], np.int32)
    get = app.getSourceSampleDouble if precision == "f64" else app.getSourceSample
    sig = np.array([[get(i, n) for n in range(n_steps)] for i in range(ns)])
    table = np.repeat(betas[:, None], 20, axis=1)
    return oracle.run(pos, mat, table, src, np.array([1] * ns, np.int32), sig, rec, n_steps,
                      precision=precision)


@pytest.mark.gpu
@pytest.mark.parametrize("double", [False, True])
def test_run_simulation_matches_oracle(gpu, double):
    n_steps = 300
    app, lab, betas = _room_app(fs=40000, n_steps=n_steps, double=double)
    app.addSource(0.15, 0.12, 0.10, 1, 1, 0)    # soft Gaussian (input type 1)
    app.addReceiver(0.2, 0.15, 0.12)
    app.addReceiver(0.05, 0.05, 0.05)
    app.runSimulation()
    prec = "f64" if double else "f32"
    ref = _oracle_responses(app, lab, betas, n_steps, prec)
    got = np.array([(app.getResponseDouble if double else app.getResponse)(r) for r in range(2)])
    assert np.array_equal(got, ref)
    assert app.getResponseSize() == 2 * n_steps
    assert app.getNumElems() == lab.size and app.getTimePerStep() > 0
    app.close()


@pytest.mark.gpu
def test_run_capture_writes_slices_and_meshes(gpu, tmp_path):
    n_steps = 60
    app, lab, betas = _room_app(fs=40000, n_steps=n_steps)
    app.addSource(0.15, 0.12, 0.10, 1, 1, 0)
    app.addReceiver(0.2, 0.15, 0.12)
    app.setCaptureDirectory(str(tmp_path))
    app.addSliceToCapture(10, 40, 0)
    app.addSliceToCapture(12, 50, 1)
    app.addSliceToCapture(7, 50, 2)
    app.addSliceToCapture(999, 50, 0)  # out of range: logged and skipped, as the reference does
    app.addMeshToCapture(25)
    app.runCapture()
    ref = _oracle_responses(app, lab, betas, n_steps, "f32")
    assert np.array_equal(app.getResponse(0), ref[0])
    files = sorted(os.listdir(tmp_path))
    assert files == ["capture_0_40_10.tga", "capture_1_50_12.tga", "capture_2_50_7.tga"]
    X, Y, Z = lab.shape[2], lab.shape[1], lab.shape[0]
    raw = (tmp_path / "capture_0_40_10.tga").read_bytes()
    assert len(raw) == 18 + 4 * X * Y and raw[2] == 2 and raw[16] == 32
    assert app.getNumberOfMeshCaptures() == 1
    pos, mat, _ = oracle.classify(lab)
    src = np.array([app.getSourceElementCoordinates(0)], np.int32)
    sig = np.array([[app.getSourceSample(0, n) for n in range(50)]])

    def field_after(n):
        return oracle.run(pos, mat, np.repeat(betas[:, None], 20, axis=1), src, np.array([1]),
                          sig[:, :n], src[:0], n, want_fields=True)[1]

    assert np.array_equal(app.getMeshCaptureAt(0), field_after(25))
    # pixel values: the oracle's slices through the reference's dB colour map
    lit = 0
    for name, step, sl, orient in (("capture_0_40_10.tga", 40, 10, 0),
                                   ("capture_1_50_12.tga", 50, 12, 1),
                                   ("capture_2_50_7.tga", 50, 7, 2)):
        p, q = oracle.capture_slice(field_after(step), pos, sl, orient)
        raw = (tmp_path / name).read_bytes()
        h, w = p.shape
        assert raw[12] | raw[13] << 8 == w and raw[14] | raw[15] << 8 == h
        got = np.frombuffer(raw[18:], np.uint8).reshape(h, w, 4)
        exp = _tga_pixels(p, q, 60.0)
        assert np.array_equal(got[..., [2, 3]], exp[..., [2, 3]])  # red and alpha
        d = np.abs(got.astype(int) - exp.astype(int))
        # green/blue: equal, up to one level where log10 of a float rounds differently
        assert d.max() <= 1 and np.count_nonzero(d) <= max(2, d.size // 1000), name
        assert np.count_nonzero((exp[..., 0] == 255) & (exp[..., 1] == 255)) > 0  # walls
        lit += np.count_nonzero((exp[..., 2] == 0) & ((exp[..., 0] > 0) | (exp[..., 1] > 0)))
    assert lit > 100  # the wave is on the captured planes
    app.close()


def _tga_pixels(p, q, capture_db):
    """(h, w, 4) BGRA bytes of App::saveBitmap (App.cpp:492-567) for a captured plane p with
    position bytes q: boundary voxels (inside, K != 6) white; elsewhere green for positive and
    blue for negative pressure, c = (log10(p^2) + dB)/dB with dB = capture_db/10, c < 0 -> 0.
    Above full scale (c > 1) the reference's float-to-byte conversion is undefined; the
    build saturates at 255."""
    p = np.asarray(p, np.float32)
    dB = np.float32(capture_db / 10.0)
    with np.errstate(divide="ignore"):
        c = (np.log10(p * p) + dB) / dB
    c = np.clip(np.where(c > 0, c, np.float32(0)), 0, 1).astype(np.float32)
    pos_sign = p > 0
    g = np.where(pos_sign, c, 0).astype(np.float32) * np.float32(255)
    b = np.where(~pos_sign, c, 0).astype(np.float32) * np.float32(255)
    out = np.zeros(p.shape + (4,), np.uint8)
    out[..., 0] = b.astype(np.uint8)
    out[..., 1] = g.astype(np.uint8)
    out[..., 3] = 255
    wall = ((q >> 7) == 1) & ((q & 0x7F) != 6)
    out[wall, 0:3] = 255
    return out


@pytest.mark.gpu
def test_update_type_srl_reclassifies_labels(gpu):
    """setUpdateType(2) after the label domain is set: the App reclassifies it for the centred
    boundaries at initializeMesh, as the reference classifies when the mesh is set up."""
    n_steps = 200
    app, lab, betas = _room_app(fs=40000, n_steps=n_steps)
    app.setUpdateType(2)
    app.addSource(0.15, 0.12, 0.10, 1, 1, 0)
    app.addReceiver(0.2, 0.15, 0.12)
    app.runSimulation()
    pos, mat, counts = oracle.classify_kow(lab)
    assert counts[2] == 0
    src = np.array([app.getSourceElementCoordinates(0)], np.int32)
    rec = np.array([app.getReceiverElementCoordinates(0)], np.int32)
    sig = np.array([[app.getSourceSample(0, n) for n in range(n_steps)]], np.float32)
    ref = oracle.run(pos, mat, np.repeat(betas[:, None], 20, axis=1), src, np.array([1], np.int32),
                     sig, rec, n_steps, literal=oracle.SCHEME_KOWALCZYK)
    assert np.array_equal(app.getResponse(0), ref[0])
    app.close()


@pytest.mark.gpu
def test_transparent_source_through_app(gpu):
    """A transparent source: the App filters its signal on the device; the responses equal the
    oracle run driven by the App's own host-side getSourceSample values."""
    n_steps = 150
    app, lab, betas = _room_app(fs=40000, n_steps=n_steps)
    h = np.zeros(200, np.float32)
    h[2::2] = (-1.0 / 3.0) * (0.95 ** np.arange(99))
    app.setGridIr(h.tolist())
    app.addSource(0.15, 0.12, 0.10, 2, 1, 0)    # transparent, Gaussian input
    app.addReceiver(0.2, 0.15, 0.12)
    app.runSimulation()
    ref = _oracle_responses(app, lab, betas, n_steps, "f32")
    assert np.array_equal(app.getResponse(0), ref[0])
    app.close()


def test_geometry_file_errors():
    app = pf.App()
    with pytest.raises(RuntimeError):
        app.initializeGeometryFromFile("/nonexistent/room.vtk")
