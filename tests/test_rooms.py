"""Seeded synthetic inputs: the generator and the room builders."""
import numpy as np

from parallelfdtd_amd import rooms


def test_splitmix64_reference_vector():
    # SplitMix64 with seed 0: first outputs of the published reference implementation
    g = rooms.SplitMix64(0)
    assert [g.next_u64() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                                0x06C45D188009454F]


def test_scenes_are_deterministic_and_shelled():
    a = rooms.scene_c2(n_steps=10, scale=8)
    b = rooms.scene_c2(n_steps=10, scale=8)
    assert np.array_equal(a.labels, b.labels) and np.array_equal(a.rec_xyz, b.rec_xyz)
    for sc in (a, rooms.scene_c3(n_steps=5, scale=16), rooms.scene_c1(n_steps=5)):
        lab = sc.labels
        assert (lab[0] > 0).all() and (lab[-1] > 0).all()
        assert (lab[:, 0] > 0).all() and (lab[:, -1] > 0).all()
        assert (lab[:, :, 0] > 0).all() and (lab[:, :, -1] > 0).all()
        for x, y, z in np.concatenate([sc.src_xyz, sc.rec_xyz]):
            assert rooms.is_air_k6(lab, x, y, z)


def test_c1_scene_matches_config():
    sc = rooms.scene_c1()
    assert sc.shape_xyz == (130, 98, 66) and sc.n_steps == 2000
    assert np.isclose(sc.signals[0].max(), 1.0) and np.argmax(sc.signals[0]) == 50
    assert sc.betas.tolist() == [np.float32(0.1)]


def test_pulses():
    g = rooms.gaussian_derivative_pulse(200, 48.0, 8.0)
    assert np.isclose(np.abs(g).max(), 1.0, atol=1e-3)
