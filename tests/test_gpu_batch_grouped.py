"""Grouped source batching (PFDTD_BATCH_GROUPED: each slot's fields contiguous, slot outermost)
through the two-level sweep, one launch for all slots. Every slot must reproduce, bit for bit,
an unbatched oracle run that holds only that slot's sources and receivers: FP32 (column and wide
blocks), FP64, FP16/BF16 storage (two-column sweep on even X, one column on odd X), odd step
counts (a single-level step per slot at the end), step mode, and slots with colliding or no
sources."""
import numpy as np
import pytest

import oracle
from parallelfdtd_amd import native as N
from parallelfdtd_amd import rooms
from parallelfdtd_amd.solver import Solver
from test_gpu_batch import PRECS, slot_oracle

pytestmark = pytest.mark.gpu


def run_grouped(sc, B, src_slot, rec_slot, prec, scale_exp=0, zchunk=0, shape=N.SWEEP_AUTO,
                literal=0, octave=0, fields_dtype=None, steps_plan=None, poll=100, levels=2):
    with Solver() as s:
        s.set_domain_labels(sc.labels)
        s.set_materials(sc.material_table(), octave=octave, literal_slot=bool(literal & 1),
                        soft_overwrite=bool(literal & 2))
        s.set_sources(sc.src_xyz, sc.src_kind, signal_f64=sc.signals,
                      signal_f32=sc.signals.astype(np.float32))
        s.set_receivers(sc.rec_xyz)
        s.set_batch(B)
        s.set_batch_layout(N.BATCH_GROUPED)
        s.set_source_slots(src_slot)
        s.set_receiver_slots(rec_slot)
        s.set_time_levels(levels, zchunk)
        s.set_sweep_shape(shape)
        s.set_poll_interval(poll)
        if prec in (N.F16, N.BF16):
            s.prepare_fields(prec, scale_exp)
        if steps_plan is None:
            done, _ = s.run(sc.n_steps, prec)
            assert done == sc.n_steps
        else:
            s.begin_steps(sc.n_steps, prec)
            for op in steps_plan:
                if op[0] == "step":
                    s.step(op[1], 1)
                else:
                    s.steps(op[1], op[2])
        resp = s.responses(np.float64 if prec == N.F64 else np.float32)
        fields = s.fields_raw(fields_dtype) if fields_dtype is not None else None
    return resp, fields


def check_slots(sc, B, src_slot, rec_slot, pname, resp, fields=None, scale_exp=0, literal=0,
                octave=0):
    pos, mat, _ = oracle.classify(sc.labels)
    dt = PRECS[pname][2]
    for slot in range(B):
        ref, P_ref, Pp_ref = slot_oracle(sc, pos, mat, src_slot, rec_slot, slot, pname,
                                         scale_exp=scale_exp, literal=literal, octave=octave,
                                         want_fields=True)
        rm = rec_slot == slot
        assert np.array_equal(resp[rm], ref.astype(dt)), f"slot {slot} responses differ"
        if fields is not None:
            P, Pp = fields
            assert P.shape == (B,) + sc.labels.shape
            conv = (lambda a: a.astype(np.float32)) if P.dtype == np.float16 else (lambda a: a)
            assert np.array_equal(conv(P[slot]), P_ref), f"slot {slot} current field differs"
            assert np.array_equal(conv(Pp[slot]), Pp_ref), f"slot {slot} past field differs"


@pytest.mark.parametrize("n_steps", [60, 61])
@pytest.mark.parametrize("B", [2, 8])
@pytest.mark.parametrize("pname", ["f32", "f64"])
def test_grouped_slots_match_unbatched_runs(gpu, B, pname, n_steps):
    sc = rooms.scene_random_small(141 + B, dims=(70, 23, 17), n_mat=3, n_boxes=4, n_src=6,
                                  n_rec=9, n_steps=n_steps)
    src_slot = np.array([(3 * i) % B for i in range(6)], np.int32)  # some slots hold none
    rec_slot = np.array([(i * 5 + 1) % B for i in range(9)], np.int32)
    dt = PRECS[pname][2]
    resp, fields = run_grouped(sc, B, src_slot, rec_slot, PRECS[pname][0], zchunk=5,
                               fields_dtype=dt)
    check_slots(sc, B, src_slot, rec_slot, pname, resp, fields)


@pytest.mark.parametrize("n_steps", [60, 62])
@pytest.mark.parametrize("pname", ["f32", "f64", "f16"])
def test_grouped_three_levels(gpu, pname, n_steps):
    """Grouped slots through the three-level sweep (asked for explicitly), with 60 steps (whole
    three-level launches) and 62 (two steps left: a two-level launch)."""
    sc = rooms.scene_random_small(151, dims=(130, 25, 15), n_mat=3, n_boxes=4, n_src=6,
                                  n_rec=9, n_steps=n_steps)
    B = 4
    src_slot = np.array([(3 * i) % B for i in range(6)], np.int32)
    rec_slot = np.array([(i * 5 + 1) % B for i in range(9)], np.int32)
    scale = 10 if pname == "f16" else 0
    dt = np.float16 if pname == "f16" else PRECS[pname][2]  # raw 16-bit fields for FP16
    resp, fields = run_grouped(sc, B, src_slot, rec_slot, PRECS[pname][0], scale_exp=scale,
                               zchunk=4, fields_dtype=dt, levels=3)
    check_slots(sc, B, src_slot, rec_slot, pname, resp, fields, scale_exp=scale)


@pytest.mark.parametrize("dims", [(130, 30, 14), (131, 30, 14)])
@pytest.mark.parametrize("pname", ["f16", "bf16"])
def test_grouped_half_storage(gpu, pname, dims):
    sc = rooms.scene_random_small(151, dims=dims, n_mat=3, n_boxes=4, n_src=5, n_rec=8,
                                  n_steps=41)
    B = 4
    src_slot = np.array([0, 1, 2, 3, 3], np.int32)
    rec_slot = np.array([3, 2, 1, 0, 0, 2, 1, 3], np.int32)
    resp, fields = run_grouped(sc, B, src_slot, rec_slot, PRECS[pname][0], scale_exp=10,
                               fields_dtype=np.float16 if pname == "f16" else None)
    check_slots(sc, B, src_slot, rec_slot, pname, resp, fields, scale_exp=10)


def test_grouped_wide_blocks_quirks_and_collisions(gpu):
    """FP32 wide blocks (4 x 4 wave tiles) on a room partly past the last block, colliding
    sources in one slot, the literal quirks, bundles of 7 steps (odd: single-level steps per
    slot in the middle of the run)."""
    sc = rooms.scene_random_small(161, dims=(250, 30, 12), n_mat=3, n_boxes=4, n_src=5, n_rec=6,
                                  n_steps=37)
    sc.src_xyz[1] = sc.src_xyz[0]
    src_slot = np.array([1, 1, 0, 2, 3], np.int32)
    rec_slot = np.array([0, 1, 1, 2, 3, 3], np.int32)
    resp, fields = run_grouped(sc, 4, src_slot, rec_slot, N.F32, shape=N.SWEEP_WIDE, literal=3,
                               octave=1, fields_dtype=np.float32, poll=7)
    check_slots(sc, 4, src_slot, rec_slot, "f32", resp, fields, literal=3, octave=1)


def test_grouped_step_mode(gpu):
    sc = rooms.scene_random_small(171, dims=(66, 20, 13), n_mat=2, n_boxes=3, n_src=4, n_rec=4,
                                  n_steps=30)
    src_slot = np.array([0, 1, 1, 0], np.int32)
    rec_slot = np.array([1, 0, 1, 0], np.int32)
    plan = [("steps", 0, 7), ("step", 7), ("steps", 8, 10), ("step", 18), ("step", 19),
            ("steps", 20, 10)]
    resp, fields = run_grouped(sc, 2, src_slot, rec_slot, N.F32, fields_dtype=np.float32,
                               steps_plan=plan)
    check_slots(sc, 2, src_slot, rec_slot, "f32", resp, fields)


def test_grouped_c4_quarter_32_slots(gpu):
    """C4 at 1/4 scale with 32 slots, 8 receivers each."""
    sc = rooms.scene_c4(n_sources=32, n_steps=40, scale=4)
    B = 32
    src_slot = np.arange(B, dtype=np.int32)
    rec_slot = np.repeat(np.arange(B, dtype=np.int32), 8)
    resp, _ = run_grouped(sc, B, src_slot, rec_slot, N.F32)
    pos, mat, _ = oracle.classify(sc.labels)
    for slot in (0, 5, 17, 31):
        ref = slot_oracle(sc, pos, mat, src_slot, rec_slot, slot, "f32")
        assert np.array_equal(resp[rec_slot == slot], ref), f"slot {slot}"
    assert np.abs(resp).max() > 0


def test_batch_layout_rejects_unknown(gpu):
    with Solver() as s:
        with pytest.raises(N.PfdtdError):
            s.set_batch_layout(2)


def test_grouped_on_recycled_device_memory(gpu):
    """The spare fields' first and last slices of every slot must be cleared even when the
    allocation comes back holding an earlier run's values: an interleaved run, then the grouped
    layout in the same context (fields reallocated), then FP64 followed by FP32 in new contexts."""
    sc = rooms.scene_random_small(181, dims=(70, 23, 17), n_mat=3, n_boxes=4, n_src=6, n_rec=9,
                                  n_steps=30)
    B = 2
    src_slot = np.array([(3 * i) % B for i in range(6)], np.int32)
    rec_slot = np.array([(i * 5 + 1) % B for i in range(9)], np.int32)
    with Solver() as s:
        s.set_domain_labels(sc.labels)
        s.set_materials(sc.material_table())
        s.set_sources(sc.src_xyz, sc.src_kind, signal_f64=sc.signals,
                      signal_f32=sc.signals.astype(np.float32))
        s.set_receivers(sc.rec_xyz)
        s.set_batch(B)
        s.set_source_slots(src_slot)
        s.set_receiver_slots(rec_slot)
        s.run(sc.n_steps, N.F32)  # interleaved: every element of the fields written
        s.set_batch_layout(N.BATCH_GROUPED)
        s.set_time_levels(2, 5)
        s.run(sc.n_steps, N.F32)
        resp = s.responses(np.float32)
        fields = s.fields_raw(np.float32)
    check_slots(sc, B, src_slot, rec_slot, "f32", resp, fields)
    for pname in ("f64", "f32"):
        resp, fields = run_grouped(sc, B, src_slot, rec_slot, PRECS[pname][0], zchunk=5,
                                   fields_dtype=PRECS[pname][2])
        check_slots(sc, B, src_slot, rec_slot, pname, resp, fields)
