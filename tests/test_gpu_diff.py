"""The difference-form FP16 storage (PFDTD_F16_DIFF: binary16 P^n and D^n = P^n - P^{n-1}, FP32
arithmetic) through the two-column sweep, against the oracle's ORC_STORE_F16_DIFF (itself pinned
by the independent numpy fixtures of tests/golden/): responses and both stored fields bit-exact,
for even and odd step counts (the odd last step takes one level of the sweep), hard, soft,
colliding and overwriting sources, sources on solid voxels and on pair/tile borders."""
import numpy as np
import pytest

import oracle
from parallelfdtd_amd import native as N
from parallelfdtd_amd import rooms
from parallelfdtd_amd.solver import Solver

pytestmark = pytest.mark.gpu


def run_diff(sc, scale, literal=0, zchunk=0, poll=100):
    pos, mat, _ = oracle.classify(sc.labels)
    sig = sc.signals.astype(np.float32)
    ref, P_ref, D_ref = oracle.run(pos, mat, sc.material_table(), sc.src_xyz, sc.src_kind, sig,
                                   sc.rec_xyz, sc.n_steps, literal=literal,
                                   storage=oracle.STORE_F16_DIFF, scale_exp=scale,
                                   want_fields=True)
    with Solver() as s:
        s.set_time_levels(2, zchunk)
        s.set_domain_labels(sc.labels)
        s.set_materials(sc.material_table(), literal_slot=bool(literal & 1),
                        soft_overwrite=bool(literal & 2))
        s.set_sources(sc.src_xyz, sc.src_kind, signal_f64=sc.signals, signal_f32=sig)
        s.set_receivers(sc.rec_xyz)
        s.set_poll_interval(poll)
        s.prepare_fields(N.F16_DIFF, scale)
        done, _ = s.run(sc.n_steps, N.F16_DIFF)
        got = s.responses(np.float32)
        P, D = s.fields_raw(np.float16)
        cap = s.capture_field()
    assert done == sc.n_steps
    assert np.abs(P_ref).max() > 0 and (sc.n_steps < 10 or np.abs(ref).max() > 0)
    assert np.array_equal(got, ref), f"responses differ: max {np.abs(got - ref).max()}"
    assert np.array_equal(P.astype(np.float32), P_ref), "P differs"
    assert np.array_equal(D.astype(np.float32), D_ref), "D differs"
    assert np.array_equal(cap, P_ref * np.float32(2.0 ** -scale))


@pytest.mark.parametrize("n_steps", [40, 41, 1, 2])
@pytest.mark.parametrize("seed", [201, 202])
def test_diff_random_rooms(gpu, seed, n_steps):
    sc = rooms.scene_random_small(seed, dims=(130, 37, 19), n_mat=3, n_boxes=5, n_src=4, n_rec=6,
                                  n_steps=n_steps)
    run_diff(sc, 9, zchunk=5)


@pytest.mark.parametrize("literal", [1, 2, 3])
def test_diff_quirk_modes(gpu, literal):
    sc = rooms.scene_random_small(203, dims=(64, 30, 16), n_mat=3, n_boxes=4, n_src=4, n_rec=4,
                                  n_steps=33)
    run_diff(sc, 8, literal=literal)


def test_diff_colliding_sources_and_odd_polls(gpu):
    """Three sources on one voxel (applied in index order, each keeping the past value), a
    hard source, and bundles of 7 steps (odd: one-level launches in the middle of the run)."""
    sc = rooms.scene_random_small(204, dims=(66, 28, 14), n_mat=2, n_boxes=2, n_src=1, n_rec=1,
                                  n_steps=37)
    rng = np.random.default_rng(3)
    sc.src_xyz = np.array([[20, 10, 7], [20, 10, 7], [20, 10, 7], [41, 17, 5], [2, 2, 2]],
                          np.int32)
    sc.src_kind = np.array([1, 1, 0, 1, 1], np.int32)
    sc.signals = rng.uniform(-1, 1, (5, sc.n_steps))
    sc.rec_xyz = np.array([[20, 10, 7], [21, 10, 7], [41, 17, 5], [2, 2, 2]], np.int32)
    run_diff(sc, 9, poll=7)


def test_diff_pair_borders(gpu):
    from test_gpu_fused import pair_border_scene
    sc = pair_border_scene(205, (250, 30, 12), n_steps=27)
    run_diff(sc, 9, zchunk=4)


def test_diff_c1_2000_steps(gpu):
    sc = rooms.scene_c1(n_steps=2000)
    run_diff(sc, 12)


def test_diff_rejects_what_it_does_not_cover(gpu):
    odd = rooms.scene_random_small(206, dims=(31, 20, 12), n_steps=10)
    with Solver() as s:
        s.load_scene(odd)
        s.prepare_fields(N.F16_DIFF, 9)
        with pytest.raises(N.PfdtdError) as e:
            s.run(10, N.F16_DIFF)
        assert e.value.code == N.E_STATE
    even = rooms.scene_random_small(207, dims=(32, 20, 12), n_steps=10)
    with Solver() as s:
        s.load_scene(even)
        with pytest.raises(N.PfdtdError) as e:
            s.begin_steps(10, N.F16_DIFF)
        assert e.value.code == N.E_ARG
        s.set_sources(np.array([[5, 5, 0]], np.int32), np.array([1], np.int32),
                      signal_f32=np.ones((1, 10), np.float32))
        with pytest.raises(N.PfdtdError) as e:
            s.run(10, N.F16_DIFF)
        assert e.value.code == N.E_STATE
