"""Coefficient tables staged in LDS by the update kernels (the (K, material) rows a domain can
address: 7 rows x (highest material byte + 1) columns per block). Domains whose material bytes
reach the top of the 0..255 range -- the largest staged tables, 56 KB in FP64 -- give the
oracle's results bit for bit in every precision and update path: single-level, two-level,
centred boundaries, source batching and the serving pool."""
import numpy as np
import pytest

import oracle
from parallelfdtd_amd import native as N
from parallelfdtd_amd import rooms, serve
from parallelfdtd_amd.solver import Solver
from test_gpu_fused import assert_same, run_both

pytestmark = pytest.mark.gpu


def high_material_scene(seed, n_mat=255, dims=(70, 40, 24), n_steps=48):
    """A random room whose boxes use the top materials (up to byte n_mat - 1)."""
    sc = rooms.scene_random_small(seed, dims=dims, n_mat=3, n_boxes=0, n_src=3, n_rec=5,
                                  n_steps=n_steps)
    rng = rooms.SplitMix64(seed + 1000)
    top = list(range(n_mat - 12, n_mat))
    rooms.add_boxes(sc.labels, rng, 14, 2, 9, top)
    assert sc.labels.max() == n_mat  # material byte n_mat - 1 is present
    r = np.random.default_rng(seed)
    sc.betas = r.uniform(0.01, 0.6, n_mat).astype(np.float32)
    return sc


@pytest.mark.parametrize("prec", [N.F32, N.F64, N.F16, N.BF16])
@pytest.mark.parametrize("levels", [1, 2])
def test_top_material_bytes_bit_exact(gpu, prec, levels):
    sc = high_material_scene(301)
    got, ref, F, F_ref = run_both(sc, prec, levels=levels, zchunk=8)
    assert np.abs(ref).max() > 0
    assert_same(got, ref, F, F_ref)


def test_literal_slot_with_material_bytes_past_the_table(gpu):
    """Literal material slot mat*20*octave at octave 0 addresses slot 0 for every byte, so a
    one-material table serves material bytes up to 254 (a 255-column staged table)."""
    sc = high_material_scene(302)
    sc.labels[sc.labels > 200] = 255  # bytes 254 on those solids
    sc.betas = np.array([0.2], np.float32)
    got, ref, F, F_ref = run_both(sc, N.F64, levels=2, zchunk=8, literal=1)
    assert_same(got, ref, F, F_ref)


def fill_without_centred_stencil(lab):
    """Make every inside voxel have an air neighbour on each axis (fill the others)."""
    for _ in range(20):
        pos, _, counts = oracle.classify_kow(lab)
        if counts[2] == 0:
            return
        inside = (pos >> 7) == 1
        air = np.pad(lab == 0, 1, constant_values=False)
        bad = np.zeros_like(inside)
        for axis in range(3):
            lo = [slice(1, -1)] * 3
            hi = [slice(1, -1)] * 3
            lo[axis] = slice(0, -2)
            hi[axis] = slice(2, None)
            bad |= inside & ~air[tuple(lo)] & ~air[tuple(hi)]
        lab[bad] = 1
    raise AssertionError("could not remove voxels without a centred stencil")


def test_centred_update_with_top_material_bytes(gpu):
    sc = high_material_scene(303, dims=(40, 30, 20), n_steps=40)
    fill_without_centred_stencil(sc.labels)
    pos, mat, _ = oracle.classify_kow(sc.labels)
    assert mat.max() >= 240
    ref = oracle.run(pos, mat, sc.material_table(), sc.src_xyz, sc.src_kind, sc.signals, sc.rec_xyz,
                     sc.n_steps, precision="f64", literal=oracle.SCHEME_KOWALCZYK)
    with Solver() as s:
        s.set_update_type(N.SRL)
        s.set_domain_labels(sc.labels)
        s.set_materials(sc.material_table())
        s.set_sources(sc.src_xyz, sc.src_kind, signal_f64=sc.signals,
                      signal_f32=sc.signals.astype(np.float32))
        s.set_receivers(sc.rec_xyz)
        s.run(sc.n_steps, N.F64)
        got = s.responses(np.float64)
    assert np.array_equal(got, ref)


def test_batched_slots_with_top_material_bytes(gpu):
    sc = high_material_scene(304, dims=(64, 24, 16), n_steps=30)
    B = 4
    src = np.repeat(sc.src_xyz[:1], B, axis=0)
    src[:, 0] = [5, 20, 40, 60]
    sig = np.stack([sc.signals[0] * (k + 1) for k in range(B)])
    rec = sc.rec_xyz[:2]
    pos, mat, _ = oracle.classify(sc.labels)
    with Solver() as s:
        s.set_domain_labels(sc.labels)
        s.set_materials(sc.material_table())
        s.set_sources(src, np.ones(B, np.int32), signal_f64=sig, signal_f32=sig.astype(np.float32))
        s.set_receivers(np.concatenate([rec] * B))
        s.set_batch(B)
        s.set_source_slots(np.arange(B, dtype=np.int32))
        s.set_receiver_slots(np.repeat(np.arange(B, dtype=np.int32), len(rec)))
        s.run(sc.n_steps, N.F32)
        got = s.responses(np.float32).reshape(B, len(rec), sc.n_steps)
    for k in range(B):
        ref = oracle.run(pos, mat, sc.material_table(), src[k:k + 1], np.ones(1, np.int32),
                         sig[k:k + 1].astype(np.float32), rec, sc.n_steps)
        assert np.array_equal(got[k], ref), f"slot {k}"


def test_pool_job_with_255_materials(gpu):
    sc = high_material_scene(305, dims=(48, 32, 20), n_steps=40)
    spec = serve.JobSpec(sc.labels, sc.betas, sc.src_xyz, sc.src_kind,
                         sc.signals.astype(np.float32), sc.rec_xyz, sc.n_steps)
    alone = serve.run_alone(spec)
    with serve.Pool(0, 1 << 28, 8) as pool:
        jid = pool.submit(spec)
        assert jid is not None
        finished = []
        for _ in range(1000):  # 8 steps per bundle: 5 bundles finish the 40 steps
            pool.advance()
            finished += pool.poll()
            if jid in finished:
                break
        assert jid in finished
        got = pool.responses(jid)
        pool.release(jid)
    assert np.array_equal(got, alone)
    pos, mat, _ = oracle.classify(sc.labels)
    ref = oracle.run(pos, mat, sc.material_table(), sc.src_xyz, sc.src_kind,
                     sc.signals.astype(np.float32), sc.rec_xyz, sc.n_steps)
    assert np.array_equal(got, ref)
