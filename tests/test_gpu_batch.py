"""Source batching on the GPU: B source configurations on one domain, one slot each, advanced
together (field layout: B values per voxel, slot innermost). Every slot must reproduce, bit for
bit, an unbatched oracle run that holds only that slot's sources and receivers.
"""
import numpy as np
import pytest

import oracle
from parallelfdtd_amd import native as N
from parallelfdtd_amd import rooms
from parallelfdtd_amd.solver import Solver

pytestmark = pytest.mark.gpu

PRECS = {"f32": (N.F32, oracle.STORE_F32, np.float32), "f64": (N.F64, oracle.STORE_F32, np.float64),
         "f16": (N.F16, oracle.STORE_F16, np.float32), "bf16": (N.BF16, oracle.STORE_BF16, np.float32)}


def slot_oracle(sc, pos, mat, src_slot, rec_slot, slot, pname, scale_exp=0, literal=0, octave=0,
                want_fields=False):
    prec, storage, _ = PRECS[pname]
    sm = np.asarray(src_slot) == slot
    rm = np.asarray(rec_slot) == slot
    sig = sc.signals if pname == "f64" else sc.signals.astype(np.float32)
    return oracle.run(pos, mat, sc.material_table(), sc.src_xyz[sm], sc.src_kind[sm], sig[sm],
                      sc.rec_xyz[rm], sc.n_steps, precision="f64" if pname == "f64" else "f32",
                      storage=storage, scale_exp=scale_exp, literal=literal, octave=octave,
                      threads=8, want_fields=want_fields)


def run_batched(sc, B, src_slot, rec_slot, prec, scale_exp=0, rows=None, zchunk=8, literal=0,
                octave=0, fields_dtype=None):
    with Solver() as s:
        s.set_domain_labels(sc.labels)
        s.set_materials(sc.material_table(), octave=octave, literal_slot=bool(literal & 1),
                        soft_overwrite=bool(literal & 2))
        s.set_sources(sc.src_xyz, sc.src_kind, signal_f64=sc.signals,
                      signal_f32=sc.signals.astype(np.float32))
        s.set_receivers(sc.rec_xyz)
        s.set_batch(B)
        s.set_source_slots(src_slot)
        s.set_receiver_slots(rec_slot)
        if rows is not None:
            s.set_update_tiling(rows, zchunk)
        if prec in (N.F16, N.BF16):
            s.prepare_fields(prec, scale_exp)
        done, _ = s.run(sc.n_steps, prec)
        assert done == sc.n_steps
        resp = s.responses(np.float64 if prec == N.F64 else np.float32)
        fields = s.fields_raw(fields_dtype) if fields_dtype is not None else None
    return resp, fields


@pytest.mark.parametrize("B", [2, 4, 8])
@pytest.mark.parametrize("pname", ["f32", "f64"])
def test_batch_slots_match_unbatched_runs(gpu, B, pname):
    """6 mixed hard/soft sources spread over B slots (some slots shared, some empty), 9
    receivers; responses and both fields of every slot equal the per-slot oracle run."""
    sc = rooms.scene_random_small(41 + B, dims=(18, 13, 11), n_mat=3, n_boxes=3, n_src=6,
                                  n_rec=9, n_steps=90)
    pos, mat, _ = oracle.classify(sc.labels)
    src_slot = np.array([(3 * i) % B for i in range(6)], np.int32)
    rec_slot = np.array([(i * 5 + 1) % B for i in range(9)], np.int32)
    dt = PRECS[pname][2]
    resp, (P, Pp) = run_batched(sc, B, src_slot, rec_slot, PRECS[pname][0], fields_dtype=dt)
    assert P.shape == sc.labels.shape + (B,)
    for slot in range(B):
        ref, P_ref, Pp_ref = slot_oracle(sc, pos, mat, src_slot, rec_slot, slot, pname,
                                         want_fields=True)
        rm = rec_slot == slot
        assert np.array_equal(resp[rm], ref.astype(dt)), f"slot {slot} responses differ"
        assert np.array_equal(P[..., slot], P_ref), f"slot {slot} current field differs"
        assert np.array_equal(Pp[..., slot], Pp_ref), f"slot {slot} past field differs"


@pytest.mark.parametrize("pname", ["f16", "bf16"])
def test_batch_half_storage(gpu, pname):
    sc = rooms.scene_random_small(51, dims=(20, 14, 12), n_mat=3, n_boxes=2, n_src=4, n_rec=6,
                                  n_steps=100)
    pos, mat, _ = oracle.classify(sc.labels)
    B = 4
    src_slot = np.array([0, 1, 2, 3], np.int32)
    rec_slot = np.array([3, 2, 1, 0, 0, 2], np.int32)
    resp, _ = run_batched(sc, B, src_slot, rec_slot, PRECS[pname][0], scale_exp=10)
    for slot in range(B):
        ref = slot_oracle(sc, pos, mat, src_slot, rec_slot, slot, pname, scale_exp=10)
        assert np.array_equal(resp[rec_slot == slot], ref), f"{pname} slot {slot}"


@pytest.mark.parametrize("literal", [1, 2, 3])
def test_batch_reference_quirks(gpu, literal):
    sc = rooms.scene_random_small(61 + literal, dims=(16, 12, 10), n_mat=3, n_boxes=3, n_src=4,
                                  n_rec=4, n_steps=70)
    pos, mat, _ = oracle.classify(sc.labels)
    octave = 1 if literal & 1 else 0
    src_slot = np.array([1, 1, 0, 1], np.int32)  # slot 1 holds colliding/overlapping sources
    sc.src_xyz[1] = sc.src_xyz[0]
    rec_slot = np.array([0, 1, 1, 0], np.int32)
    resp, _ = run_batched(sc, 2, src_slot, rec_slot, N.F32, literal=literal, octave=octave)
    for slot in range(2):
        ref = slot_oracle(sc, pos, mat, src_slot, rec_slot, slot, "f32", literal=literal,
                          octave=octave)
        assert np.array_equal(resp[rec_slot == slot], ref), f"quirks {literal} slot {slot}"


@pytest.mark.parametrize("rows,zchunk", [(1, 3), (2, 8), (4, 5)])
def test_batch_tiling_invariance(gpu, rows, zchunk):
    sc = rooms.scene_random_small(71, dims=(37, 21, 15), n_mat=3, n_boxes=3, n_src=3, n_rec=5,
                                  n_steps=60)
    pos, mat, _ = oracle.classify(sc.labels)
    B = 16  # 16 x-columns per 256-thread block: several blocks along x
    src_slot = np.array([0, 7, 15], np.int32)
    rec_slot = np.array([0, 7, 15, 15, 3], np.int32)
    resp, _ = run_batched(sc, B, src_slot, rec_slot, N.F32, rows=rows, zchunk=zchunk)
    for slot in (0, 3, 7, 15):
        ref = slot_oracle(sc, pos, mat, src_slot, rec_slot, slot, "f32")
        assert np.array_equal(resp[rec_slot == slot], ref), f"slot {slot}"


def test_batch_c4_sources(gpu):
    """512x384x256 multi-source room at 1/4 scale: 32 sources in 32 slots, 8 receivers each."""
    sc = rooms.scene_c4(n_sources=32, n_steps=120, scale=4)
    pos, mat, _ = oracle.classify(sc.labels)
    src_slot = np.arange(32, dtype=np.int32)
    rec_slot = np.repeat(np.arange(32, dtype=np.int32), 8)
    resp, _ = run_batched(sc, 32, src_slot, rec_slot, N.F32)
    assert np.abs(resp).max() > 0
    for slot in (0, 13, 31):
        ref = slot_oracle(sc, pos, mat, src_slot, rec_slot, slot, "f32")
        assert np.array_equal(resp[rec_slot == slot], ref), f"slot {slot}"


def test_batch_one_and_switch_back(gpu):
    """set_batch(1) is the unbatched path; a context switched 4 -> 1 runs unbatched again."""
    sc = rooms.scene_random_small(81, dims=(15, 12, 10), n_src=2, n_rec=3, n_steps=50)
    pos, mat, _ = oracle.classify(sc.labels)
    ref = oracle.run(pos, mat, sc.material_table(), sc.src_xyz, sc.src_kind,
                     sc.signals.astype(np.float32), sc.rec_xyz, sc.n_steps, threads=4)
    with Solver() as s:
        s.load_scene(sc)
        s.set_batch(4)
        s.set_source_slots([2, 2])
        s.set_receiver_slots([2, 2, 2])
        s.run(sc.n_steps, N.F32)
        assert np.array_equal(s.responses(np.float32), ref)
        with pytest.raises(N.PfdtdError) as ei:
            s.capture_field()
        assert ei.value.code == N.E_STATE
        s.set_batch(1)  # slots reset to 0
        s.run(sc.n_steps, N.F32)
        assert np.array_equal(s.responses(np.float32), ref)
        assert np.array_equal(s.capture_field(), oracle.run(
            pos, mat, sc.material_table(), sc.src_xyz, sc.src_kind, sc.signals.astype(np.float32),
            sc.rec_xyz, sc.n_steps, threads=4, want_fields=True)[1])


def test_batch_argument_checks(gpu):
    sc = rooms.scene_random_small(82, dims=(12, 10, 9), n_src=2, n_rec=2, n_steps=10)
    with Solver() as s:
        s.load_scene(sc)
        for bad in (0, 3, 12, 512, -4):
            with pytest.raises(N.PfdtdError) as ei:
                s.set_batch(bad)
            assert ei.value.code == N.E_ARG
        s.set_batch(2)
        with pytest.raises(N.PfdtdError):
            s.set_source_slots([0, 2])      # outside the batch
        with pytest.raises(N.PfdtdError):
            s.set_source_slots([0])         # wrong count
        with pytest.raises(N.PfdtdError):
            s.set_receiver_slots([-1, 0])


def test_autotune_sweep_small_room(gpu):
    """The tuner times single-level candidates for every B, the two- and three-level paths for
    B = 1 and the grouped layout through both for B > 1, and chooses the fastest per source."""
    from parallelfdtd_amd import autotune
    scene = rooms.scene_c4(n_sources=8, n_steps=6, scale=8)
    rep = autotune.sweep(scene, batches=(1, 8), rows=(2,), zchunks=(8,), steps=6, warmup=1,
                         log=lambda m: None)
    levels = sorted((c["B"], c["levels"], c["shape"] or "") for c in rep["candidates"])
    multi = [(2, "column")] * 3 + [(2, "wide")] * 3 + [(3, "square")] * 2
    assert levels == ([(1, 1, "")] + [(1,) + t for t in multi] + [(8, 1, "")]
                      + [(8,) + t for t in multi])
    assert set(rep["best"]) == {"1", "8"} and set(rep["best_single_level"]) == {"1", "8"}
    assert rep["best_single_level"]["1"]["levels"] == 1
    top = max(rep["candidates"], key=lambda c: c["per_source_gvox"])
    assert rep["choice"]["B"] == top["B"] and rep["choice"]["levels"] == top["levels"]
