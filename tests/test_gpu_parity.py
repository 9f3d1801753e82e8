"""GPU parity: libpfdtd.so (through the C ABI) against the CPU oracle on identical inputs.

Bar: position/material bytes and every integer/index quantity bit-exact; FP32 receiver
series within 1e-5 relative L2 of the oracle (the kernels are built to be bit-identical, and
the random-room tests assert bit equality outright); FP64 likewise.
"""
import math

import numpy as np
import pytest

import oracle
from parallelfdtd_amd import native as N
from parallelfdtd_amd import rooms
from parallelfdtd_amd.solver import Solver

pytestmark = pytest.mark.gpu

REL_L2_F32 = 1e-5


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def oracle_for(scene, pos, mat, precision="f32", literal=0, storage=0, scale_exp=0, n_steps=None,
               want_fields=False):
    sig = scene.signals if precision == "f64" else scene.signals.astype(np.float32)
    return oracle.run(pos, mat, scene.material_table(), scene.src_xyz, scene.src_kind, sig,
                      scene.rec_xyz, n_steps or scene.n_steps, precision=precision,
                      literal=literal, storage=storage, scale_exp=scale_exp, threads=8,
                      want_fields=want_fields)


def test_c1_classification_bit_exact(gpu):
    sc = rooms.scene_c1(n_steps=10)
    pos_o, mat_o, counts = oracle.classify(sc.labels)
    with Solver() as s:
        s.set_domain_labels(sc.labels)
        pos, mat = s.domain()
        air, bnd = s.counts()
    assert np.array_equal(pos, pos_o)
    assert np.array_equal(mat, mat_o)
    assert (air, bnd) == (int(counts[0]), int(counts[1])) == (734328, 52104)


def test_c1_fp32_2000_steps(gpu):
    """Config 1: 130x98x66 box, beta 0.1, soft Gaussian source, 2000 steps, sample by sample."""
    sc = rooms.scene_c1(n_steps=2000)
    pos, mat, _ = oracle.classify(sc.labels)
    ref = oracle_for(sc, pos, mat)
    with Solver() as s:
        s.load_scene(sc)
        done, _ = s.run(sc.n_steps, N.F32)
        got = s.responses(np.float32)
    assert done == sc.n_steps
    assert np.abs(ref).max() > 1e-3
    err = rel_l2(got, ref)
    print(f"C1 FP32 rel L2 = {err:.3e}, bit-exact = {np.array_equal(got, ref)}", flush=True)
    assert err <= REL_L2_F32
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("seed", [11, 12, 13, 14])
@pytest.mark.parametrize("literal", [0, 1, 2, 3])
def test_random_rooms_bit_exact(gpu, seed, literal):
    """Mixed hard/soft sources (some on solid voxels), 3 materials, both quirk switches."""
    sc = rooms.scene_random_small(seed, dims=(17, 13, 11), n_mat=3, n_boxes=3, n_src=3,
                                  n_rec=4, n_steps=80)
    pos, mat, _ = oracle.classify(sc.labels)
    octave = 1 if literal & 1 else 0  # literal slot at octave 1 = m*20, a valid slot
    for prec, pname in ((N.F32, "f32"), (N.F64, "f64")):
        ref, P_ref, Pp_ref = oracle.run(
            pos, mat, sc.material_table(), sc.src_xyz, sc.src_kind,
            sc.signals if pname == "f64" else sc.signals.astype(np.float32), sc.rec_xyz,
            sc.n_steps, precision=pname, octave=octave, literal=literal, want_fields=True)
        with Solver() as s:
            s.set_domain_labels(sc.labels)
            s.set_materials(sc.material_table(), octave=octave, literal_slot=bool(literal & 1),
                            soft_overwrite=bool(literal & 2))
            s.set_sources(sc.src_xyz, sc.src_kind, signal_f64=sc.signals,
                          signal_f32=sc.signals.astype(np.float32))
            s.set_receivers(sc.rec_xyz)
            s.run(sc.n_steps, prec)
            got = s.responses(np.float64 if pname == "f64" else np.float32)
            P, Pp = s.fields_raw(np.float64 if pname == "f64" else np.float32)
        assert np.array_equal(got, ref), f"{pname} responses differ, rel L2 {rel_l2(got, ref):.3e}"
        assert np.array_equal(P, P_ref) and np.array_equal(Pp, Pp_ref), f"{pname} fields differ"


@pytest.mark.parametrize("precision,storage,dtype", [(N.F16, oracle.STORE_F16, np.float16),
                                                     (N.BF16, oracle.STORE_BF16, np.uint16)])
def test_half_storage_bit_exact(gpu, precision, storage, dtype):
    sc = rooms.scene_random_small(21, dims=(20, 16, 12), n_mat=3, n_boxes=2, n_src=2, n_rec=3,
                                  n_steps=100)
    pos, mat, _ = oracle.classify(sc.labels)
    scale = 10
    ref = oracle_for(sc, pos, mat, storage=storage, scale_exp=scale)
    with Solver() as s:
        s.load_scene(sc)
        s.prepare_fields(precision, scale)
        s.run(sc.n_steps, precision)
        got = s.responses(np.float32)
    assert np.array_equal(got, ref), f"rel L2 {rel_l2(got, ref):.3e}"


def test_capture_slices_and_field(gpu):
    sc = rooms.scene_random_small(31, dims=(19, 14, 12), n_steps=40)
    pos, mat, _ = oracle.classify(sc.labels)
    _, P_ref, _ = oracle_for(sc, pos, mat, want_fields=True)
    with Solver() as s:
        s.load_scene(sc)
        s.run(sc.n_steps, N.F32)
        full = s.capture_field()
        assert np.array_equal(full, P_ref)
        Z, Y, X = sc.labels.shape
        for orient, lim in ((0, Z), (1, Y), (2, X)):
            for sl in (0, lim // 2, lim - 1):
                p, q = s.capture_slice(sl, orient)
                if orient == 0:
                    ep, eq = P_ref[sl], pos[sl]
                elif orient == 1:
                    ep, eq = P_ref[:, sl, :], pos[:, sl, :]
                else:
                    ep, eq = P_ref[:, :, sl], pos[:, :, sl]
                assert np.array_equal(p, ep) and np.array_equal(q, eq), (orient, sl)
        with pytest.raises(N.PfdtdError):
            s.capture_slice(Z, 0)


def test_codes_to_bilbao(gpu):
    rng = np.random.default_rng(5)
    codes = rng.integers(0, 28, size=(6, 5, 7), dtype=np.uint8)
    mats = rng.integers(0, 4, size=codes.shape, dtype=np.uint8)
    exp_pos = np.zeros_like(codes)
    exp_mat = mats.copy()
    lib = oracle.lib()
    for i, k in enumerate(codes.ravel()):
        m = np.zeros(1, np.uint8)
        m[0] = mats.ravel()[i]
        exp_pos.ravel()[i] = lib.orc_to_bilbao(int(k), m.ctypes.data)
        exp_mat.ravel()[i] = m[0]
    with Solver() as s:
        s.set_domain_codes(codes, mats)
        pos, mat = s.domain()
    assert np.array_equal(pos, exp_pos) and np.array_equal(mat, exp_mat)


def test_domain_preconditions(gpu):
    sc = rooms.scene_random_small(41, n_steps=5)
    pos, mat, _ = oracle.classify(sc.labels)
    bad = pos.copy()
    bad[0, 3, 3] = 0x86  # inside voxel on the outer layer
    with Solver() as s:
        s.set_domain(bad, mat)
        s.set_materials(sc.material_table())
        with pytest.raises(N.PfdtdError) as e:
            s.run(3, N.F32)
        assert e.value.code == N.E_DOMAIN
        bad2 = pos.copy()
        bad2[3, 3, 3] = 0x87  # K = 7 cannot occur
        s.set_domain(bad2, mat)
        with pytest.raises(N.PfdtdError) as e:
            s.run(3, N.F32)
        assert e.value.code == N.E_DOMAIN
        s.set_domain(pos, np.where(pos > 0, 9, 0).astype(np.uint8))  # material 9 of 3
        with pytest.raises(N.PfdtdError) as e:
            s.run(3, N.F32)
        assert e.value.code == N.E_DOMAIN


def test_step_mode_matches_run(gpu):
    sc = rooms.scene_random_small(51, dims=(15, 12, 10), n_steps=50)
    with Solver() as s:
        s.load_scene(sc)
        s.run(sc.n_steps, N.F32)
        ref = s.responses(np.float32)
        s.begin_steps(sc.n_steps, N.F32)
        for n in range(sc.n_steps):
            s.step(n, 1)
        got = s.responses(np.float32)
    assert np.array_equal(got, ref)


def test_interrupt_leaves_zeros(gpu):
    sc = rooms.scene_random_small(61, dims=(15, 12, 10), n_steps=60)
    calls = []
    with Solver() as s:
        s.load_scene(sc)
        s.set_poll_interval(10)
        full_done, _ = s.run(sc.n_steps, N.F32)
        full = s.responses(np.float32)
        done, _ = s.run(sc.n_steps, N.F32, interrupt=lambda: len(calls) >= 3 or calls.append(1))
        part = s.responses(np.float32)
    assert full_done == sc.n_steps and done == 30
    assert np.array_equal(part[:, :30], full[:, :30])
    assert not part[:, 30:].any()


@pytest.mark.parametrize("rows,zchunk", [(1, 8), (2, 3), (4, 1), (1, 1000)])
@pytest.mark.parametrize("prec", [N.F32, N.F64, N.F16])
def test_tiling_invariance(gpu, rows, zchunk, prec):
    """Every update tiling gives the oracle's bits (the analogue of the reference's
    1/2/5-partition bit-identity test, CudaMeshTest.cpp:472-575). Y is not a multiple of
    any tile height, and Z is split into several z-chunks."""
    sc = rooms.scene_random_small(71, dims=(300, 37, 23), n_mat=4, n_boxes=4, n_src=2, n_rec=5,
                                  n_steps=64)
    pos, mat, _ = oracle.classify(sc.labels)
    pname = "f64" if prec == N.F64 else "f32"
    storage = oracle.STORE_F16 if prec == N.F16 else 0
    scale = 8 if prec == N.F16 else 0
    ref, P_ref, Pp_ref = oracle_for(sc, pos, mat, precision=pname, storage=storage,
                                    scale_exp=scale, want_fields=True)
    with Solver() as s:
        s.load_scene(sc)
        s.set_time_levels(1)
        s.set_update_tiling(rows, zchunk)
        if prec == N.F16:
            s.prepare_fields(N.F16, scale)
        s.run(sc.n_steps, prec)
        got = s.responses(np.float64 if pname == "f64" else np.float32)
        if prec != N.F16:
            P, Pp = s.fields_raw(np.float64 if pname == "f64" else np.float32)
            assert np.array_equal(P, P_ref) and np.array_equal(Pp, Pp_ref)
    assert np.array_equal(got, ref)


def _step_sequence(sc, prec, plan, scale=0):
    """Run `plan` (a list of ("step", n, direction) / ("steps", first, count)) after
    begin_steps; return responses and raw fields."""
    raw = np.float32 if prec == N.F32 else np.float16
    with Solver() as s:
        s.load_scene(sc)
        if scale:
            s.prepare_fields(prec, scale)
        s.begin_steps(sc.n_steps, prec)
        for op in plan:
            if op[0] == "step":
                s.step(op[1], op[2])
            else:
                s.steps(op[1], op[2])
        return s.responses(np.float32), s.fields_raw(raw)


@pytest.mark.parametrize("prec,scale", [(N.F32, 0), (N.F16, 9)])
def test_steps_equal_single_step_calls(gpu, prec, scale):
    """pfdtd_steps over segments of odd and even lengths (three- and two-level sweeps inside for
    FP32, two-level for FP16) gives the fields and responses of the same steps taken one
    pfdtd_step call at a time, including after a backward step (the flip rule)."""
    sc = rooms.scene_random_small(52, dims=(130, 40, 21), n_mat=3, n_boxes=4, n_src=3, n_rec=5,
                                  n_steps=60)
    one = [("step", n, 1) for n in range(25)] + [("step", 25, -1)] + \
          [("step", n, 1) for n in range(26, 60)]
    seg = [("steps", 0, 7), ("steps", 7, 2), ("steps", 9, 1), ("steps", 10, 15), ("step", 25, -1),
           ("steps", 26, 13), ("steps", 39, 21)]
    r1, f1 = _step_sequence(sc, prec, one, scale)
    r2, f2 = _step_sequence(sc, prec, seg, scale)
    assert np.abs(r1).max() > 0
    assert np.array_equal(r1, r2)
    assert np.array_equal(f1[0], f2[0]) and np.array_equal(f1[1], f2[1])


def test_steps_match_run_and_reject_overrun(gpu):
    sc = rooms.scene_random_small(53, dims=(70, 30, 17), n_steps=41)
    with Solver() as s:
        s.load_scene(sc)
        s.run(sc.n_steps, N.F32)
        ref = s.responses(np.float32)
        s.begin_steps(sc.n_steps, N.F32)
        s.steps(0, sc.n_steps)
        assert np.array_equal(s.responses(np.float32), ref)
        s.begin_steps(sc.n_steps, N.F32)
        with pytest.raises(N.PfdtdError):
            s.steps(30, 12)
        with pytest.raises(N.PfdtdError):
            s.steps(-1, 2)


def _app_step_calls(n_forward, back, n_steps):
    """(step, direction) per pfdtd_step call as App::executeStep issues them
    (App.cpp:407-436): current_step_ += direction after each call; invertTime flips the
    direction (App.cpp:403-405)."""
    steps, dirs = [], []
    cur, d = 0, 1
    for count, direction in ((n_forward, 1), (back, -1), (n_steps, 1)):
        d = direction
        for _ in range(count):
            if not 0 <= cur < n_steps:
                break
            steps.append(cur)
            dirs.append(d)
            cur += d
    return steps, dirs


@pytest.mark.parametrize("prec,storage,scale", [(N.F32, oracle.STORE_F32, 0),
                                                (N.F16, oracle.STORE_F16, 9)])
def test_step_directions_match_oracle(gpu, prec, storage, scale):
    """Forward, backward, forward pfdtd_step sequences against the oracle's single-step mode
    (orc_step_calls_f): the fields swap only while the direction is unchanged
    (kernels3d.cu:462-465), and each call writes response sample `step`."""
    sc = rooms.scene_random_small(54, dims=(70, 26, 17), n_mat=3, n_boxes=4, n_src=3, n_rec=5,
                                  n_steps=60)
    pos, mat, _ = oracle.classify(sc.labels)
    steps, dirs = _app_step_calls(25, 6, 60)
    steps2, dirs2 = _app_step_calls(10, 1, 60)  # a single backward step
    for st, dr in ((steps, dirs), (steps2, dirs2)):
        ref, Q, Qp = oracle.step_calls(pos, mat, sc.material_table(), sc.src_xyz, sc.src_kind,
                                       sc.signals, sc.rec_xyz, sc.n_steps, st, dr,
                                       storage=storage, scale_exp=scale)
        raw = np.float32 if prec == N.F32 else np.float16
        with Solver() as s:
            s.load_scene(sc)
            if scale:
                s.prepare_fields(prec, scale)
            s.begin_steps(sc.n_steps, prec)
            for n, d in zip(st, dr):
                s.step(n, d)
            got = s.responses(np.float32)
            P, Pp = s.fields_raw(raw)
        assert np.abs(ref).max() > 0
        assert np.array_equal(got, ref)
        assert np.array_equal(P.astype(np.float32), Q) and np.array_equal(Pp.astype(np.float32), Qp)


@pytest.mark.parametrize("prec,storage,scale", [(N.F32, oracle.STORE_F32, 0),
                                                (N.F16, oracle.STORE_F16, 7)])
def test_capture_slice_matches_oracle_capture(gpu, prec, storage, scale):
    """pfdtd_capture_slice against the oracle's captureSliceFast restatement
    (orc_capture_slice_f, visualizationUtils.cu:130-254, :500-527) on the oracle's field:
    every slice of every orientation, and the out-of-range slices it refuses."""
    sc = rooms.scene_random_small(32, dims=(23, 17, 13), n_steps=30)
    pos, mat, _ = oracle.classify(sc.labels)
    _, P_ref, _ = oracle_for(sc, pos, mat, storage=storage, scale_exp=scale, want_fields=True)
    P_ref = P_ref * np.float32(2.0 ** -scale)
    Z, Y, X = sc.labels.shape
    with Solver() as s:
        s.load_scene(sc)
        if scale:
            s.prepare_fields(prec, scale)
        s.run(sc.n_steps, prec)
        for orient, lim in ((0, Z), (1, Y), (2, X)):
            for sl in range(lim):
                p, q = s.capture_slice(sl, orient)
                ep, eq = oracle.capture_slice(P_ref, pos, sl, orient)
                assert np.array_equal(p, ep) and np.array_equal(q, eq), (orient, sl)
            assert oracle.capture_slice(P_ref, pos, lim, orient) is None
            with pytest.raises(N.PfdtdError):
                s.capture_slice(lim, orient)
            with pytest.raises(N.PfdtdError):
                s.capture_slice(-1, orient)
