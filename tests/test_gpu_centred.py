"""GPU parity of update type SRL (centred Kowalczyk boundaries) against the oracle's centred
sweep: classification bytes bit-exact, responses and fields bit-exact in FP32/FP64 and with
FP16/BF16 storage; update type SHARED equals SRL_FORWARD."""
import numpy as np
import pytest

import oracle
from parallelfdtd_amd import native as N
from parallelfdtd_amd import rooms
from parallelfdtd_amd.solver import Solver

pytestmark = pytest.mark.gpu

PREC = {N.F32: ("f32", oracle.STORE_F32, 0, np.float32), N.F64: ("f64", oracle.STORE_F32, 0, np.float64),
        N.F16: ("f32", oracle.STORE_F16, 10, np.float32), N.BF16: ("f32", oracle.STORE_BF16, 10, np.float32)}


def centred_room(seed, dims=(21, 17, 13), n_src=3, n_rec=4, n_steps=70):
    """Random room whose inside voxels all have a centred stencil (voxels with no air
    neighbour on some axis are filled)."""
    sc = rooms.scene_random_small(seed, dims=dims, n_mat=3, n_boxes=3, n_src=n_src, n_rec=n_rec,
                                  n_steps=n_steps)
    sc.labels = fill_centred(sc.labels.copy())
    return sc


def fill_centred(lab):
    """Fill the inside voxels that have no air neighbour on some axis until none is left."""
    for _ in range(20):
        pos, _, counts = oracle.classify_kow(lab)
        if counts[2] == 0:
            break
        Z, Y, X = lab.shape
        inside = (pos >> 7) == 1
        bad = np.zeros_like(inside)
        air = lab == 0
        pad = np.pad(air, 1, constant_values=False)
        for axis in range(3):
            lo = [slice(1, -1)] * 3
            hi = [slice(1, -1)] * 3
            lo[axis] = slice(0, -2)
            hi[axis] = slice(2, None)
            bad |= inside & ~pad[tuple(lo)] & ~pad[tuple(hi)]
        lab[bad] = 1
    return lab


def run_gpu(sc, prec, scale=0, update_type=N.SRL, want_fields=False, step_mode=False):
    with Solver() as s:
        s.set_update_type(update_type)
        s.set_domain_labels(sc.labels)
        s.set_materials(sc.material_table())
        s.set_sources(sc.src_xyz, sc.src_kind, signal_f64=sc.signals,
                      signal_f32=sc.signals.astype(np.float32))
        s.set_receivers(sc.rec_xyz)
        if scale:
            s.prepare_fields(prec, scale)
        if step_mode:
            s.begin_steps(sc.n_steps, prec)
            for n in range(sc.n_steps):
                s.step(n)
        else:
            s.run(sc.n_steps, prec)
        got = s.responses(np.float64 if prec == N.F64 else np.float32)
        F = s.fields_raw(np.float64 if prec == N.F64 else np.float32) if want_fields else None
        dom = s.domain()
    return got, F, dom


def oracle_centred(sc, prec, want_fields=False):
    pname, storage, scale, _ = PREC[prec]
    pos, mat, _ = oracle.classify_kow(sc.labels)
    sig = sc.signals if pname == "f64" else sc.signals.astype(np.float32)
    return oracle.run(pos, mat, sc.material_table(), sc.src_xyz, sc.src_kind, sig, sc.rec_xyz,
                      sc.n_steps, precision=pname, storage=storage, scale_exp=scale,
                      literal=oracle.SCHEME_KOWALCZYK, threads=8, want_fields=want_fields)


@pytest.mark.parametrize("seed", [101, 102, 103])
def test_centred_classification_bit_exact(gpu, seed):
    sc = centred_room(seed)
    pos_o, mat_o, _ = oracle.classify_kow(sc.labels)
    _, _, (pos, mat) = run_gpu(sc, N.F32)
    assert np.array_equal(pos, pos_o) and np.array_equal(mat, mat_o)


@pytest.mark.parametrize("seed", [101, 102, 103])
@pytest.mark.parametrize("prec", [N.F32, N.F64])
def test_centred_bit_exact(gpu, seed, prec):
    sc = centred_room(seed)
    ref, P_ref, Pp_ref = oracle_centred(sc, prec, want_fields=True)
    got, (P, Pp), _ = run_gpu(sc, prec, want_fields=True)
    assert np.abs(ref).max() > 1e-3
    assert np.array_equal(got, ref), f"max diff {np.abs(got - ref).max()}"
    assert np.array_equal(P, P_ref) and np.array_equal(Pp, Pp_ref)


@pytest.mark.parametrize("prec", [N.F16, N.BF16])
def test_centred_half_storage(gpu, prec):
    sc = centred_room(104, dims=(24, 18, 14), n_steps=90)
    ref = oracle_centred(sc, prec)
    got, _, _ = run_gpu(sc, prec, scale=PREC[prec][2])
    assert np.array_equal(got, ref)


def test_centred_codes_path(gpu):
    """Arbitrary voxelizer codes 0..27 (any DIR/SIGN combination) through toKowalczyk."""
    rng = np.random.default_rng(9)
    Z, Y, X = 12, 15, 19
    codes = rng.integers(0, 28, size=(Z, Y, X), dtype=np.uint8)
    codes[[0, -1], :, :] = 0
    codes[:, [0, -1], :] = 0
    codes[:, :, [0, -1]] = 0
    mats = rng.integers(0, 3, size=codes.shape, dtype=np.uint8)
    L = oracle.lib()
    pos = np.array([L.orc_to_kowalczyk(int(k), None) for k in codes.ravel()], np.uint8).reshape(codes.shape)
    mat = np.where(codes == 0, 0, mats).astype(np.uint8)
    table = np.repeat(np.array([0.2, 0.05, 0.4], np.float32)[:, None], 20, axis=1)
    src = np.array([[5, 7, 6], [9, 3, 4]], np.int32)
    rec = np.array([[5, 7, 6], [10, 8, 8], [3, 12, 2]], np.int32)
    sig = rng.uniform(-1, 1, (2, 50))
    ref = oracle.run(pos, mat, table, src, np.array([1, 0], np.int32), sig.astype(np.float32), rec,
                     50, literal=oracle.SCHEME_KOWALCZYK)
    with Solver() as s:
        s.set_update_type(N.SRL)
        s.set_domain_codes(codes, mats)
        gpos, gmat = s.domain()
        s.set_materials(table)
        s.set_sources(src, np.array([1, 0], np.int32), signal_f32=sig.astype(np.float32))
        s.set_receivers(rec)
        s.run(50, N.F32)
        got = s.responses(np.float32)
    assert np.array_equal(gpos, pos) and np.array_equal(gmat, mat)
    assert np.array_equal(got, ref)


def test_centred_step_mode_equals_run(gpu):
    sc = centred_room(105, n_steps=40)
    a, _, _ = run_gpu(sc, N.F32)
    b, _, _ = run_gpu(sc, N.F32, step_mode=True)
    assert np.array_equal(a, b)


def test_shared_type_is_the_forward_update(gpu):
    sc = rooms.scene_random_small(106, dims=(20, 16, 12), n_steps=50)
    a, _, _ = run_gpu(sc, N.F32, update_type=N.SRL_FORWARD)
    b, _, _ = run_gpu(sc, N.F32, update_type=N.SHARED)
    pos, mat, _ = oracle.classify(sc.labels)
    ref = oracle.run(pos, mat, sc.material_table(), sc.src_xyz, sc.src_kind,
                     sc.signals.astype(np.float32), sc.rec_xyz, sc.n_steps)
    assert np.array_equal(a, ref) and np.array_equal(b, ref)


def test_centred_rejects_degenerate_and_stale_domains(gpu):
    lab = rooms.box_room(12, 10, 9)
    lab[4, 4, 1:11] = 1
    lab[4, 4, 6] = 0  # one-voxel slit: no air neighbour along z or y there
    with Solver() as s:
        s.set_update_type(N.SRL)
        s.set_domain_labels(lab)
        s.set_materials(np.full((1, 20), 0.1, np.float32))
        with pytest.raises(N.PfdtdError) as ei:
            s.run(5, N.F32)
        assert ei.value.code == N.E_DOMAIN
        # switching the boundary family drops the classified domain
        ok = rooms.box_room(12, 10, 9)
        s.set_domain_labels(ok)
        s.set_update_type(N.SRL_FORWARD)
        with pytest.raises(N.PfdtdError) as ei:
            s.run(5, N.F32)
        assert ei.value.code == N.E_STATE
        s.set_domain_labels(ok)
        s.run(5, N.F32)


def centred_border_room(seed, dims=(185, 54, 22), n_steps=36):
    """A room crossing several two-level wave tiles (60 columns, 6/8-row tiles) with single solid
    voxels next to the tile borders, sources spread over it; filled for centred stencils."""
    sc = rooms.scene_random_small(seed, dims=dims, n_mat=3, n_boxes=2, n_src=1, n_rec=1,
                                  n_steps=n_steps)
    X, Y, Z = dims
    xs = [x for t in range(1, (X + 59) // 60) for x in (60 * t - 2, 60 * t + 1) if 2 < x < X - 3]
    ys = [y for ty in (4, 6, 8) for m in range(ty, Y, 2 * ty) for y in (m - 2, m + 1) if 2 < y < Y - 3]
    for k, (x, y) in enumerate((x, y) for x in xs for y in ys):
        sc.labels[2 + (5 * k) % (Z - 4), y, x] = 2 + k % 2
    sc.labels = fill_centred(sc.labels)
    rng = np.random.default_rng(seed)
    n = 24
    sc.src_xyz = np.stack([rng.integers(2, X - 2, n), rng.integers(2, Y - 2, n),
                           rng.integers(2, Z - 2, n)], axis=1).astype(np.int32)
    sc.src_kind = np.ones(n, np.int32)
    sc.signals = rng.uniform(-1, 1, (n, n_steps))
    sc.rec_xyz = sc.src_xyz[::2].copy()
    return sc


@pytest.mark.parametrize("levels", [1, 2, 3])
@pytest.mark.parametrize("prec", [N.F32, N.F64, N.F16, N.BF16])
def test_centred_time_levels_bit_exact(gpu, prec, levels):
    """The centred update through the two-level sweep (default; also when three levels are
    asked for) and single steps, on a room several wave tiles wide with boundary voxels next to
    the tile borders."""
    sc = centred_border_room(105)
    pname, storage, scale, _ = PREC[prec]
    pos, mat, _ = oracle.classify_kow(sc.labels)
    sig = sc.signals if pname == "f64" else sc.signals.astype(np.float32)
    ref, P_ref, Pp_ref = oracle.run(pos, mat, sc.material_table(), sc.src_xyz, sc.src_kind, sig,
                                    sc.rec_xyz, sc.n_steps, precision=pname, storage=storage,
                                    scale_exp=scale, literal=oracle.SCHEME_KOWALCZYK, threads=8,
                                    want_fields=True)
    raw = {N.F32: np.float32, N.F64: np.float64, N.F16: np.float16, N.BF16: np.uint16}[prec]
    with Solver() as s:
        s.set_time_levels(levels, 5)
        s.set_update_type(N.SRL)
        s.set_domain_labels(sc.labels)
        s.set_materials(sc.material_table())
        s.set_sources(sc.src_xyz, sc.src_kind, signal_f64=sc.signals,
                      signal_f32=sc.signals.astype(np.float32))
        s.set_receivers(sc.rec_xyz)
        if scale:
            s.prepare_fields(prec, scale)
        s.run(sc.n_steps, prec)
        got = s.responses(np.float64 if prec == N.F64 else np.float32)
        P, Pp = s.fields_raw(raw)
    if raw == np.float16:
        P, Pp = P.astype(np.float32), Pp.astype(np.float32)
    elif raw == np.uint16:
        P, Pp = [(h.astype(np.uint32) << 16).view(np.float32) for h in (P, Pp)]
    assert np.abs(ref).max() > 1e-3
    assert np.array_equal(got, ref), f"max diff {np.abs(got - ref).max()}"
    assert np.array_equal(P, P_ref) and np.array_equal(Pp, Pp_ref)


@pytest.mark.parametrize("prec", [N.F32, N.F64])
def test_centred_wide_room_bit_exact(gpu, prec):
    """A room 961 columns wide (17 wave tiles of 60): FP32 takes blocks of 2 x 2 wave tiles there,
    the last block column with one tile past the grid; FP64 keeps the column of 1 x 4."""
    sc = centred_border_room(107, dims=(961, 30, 10), n_steps=24)
    pname, storage, scale, _ = PREC[prec]
    pos, mat, _ = oracle.classify_kow(sc.labels)
    sig = sc.signals if pname == "f64" else sc.signals.astype(np.float32)
    ref, P_ref, Pp_ref = oracle.run(pos, mat, sc.material_table(), sc.src_xyz, sc.src_kind, sig,
                                    sc.rec_xyz, sc.n_steps, precision=pname, storage=storage,
                                    scale_exp=scale, literal=oracle.SCHEME_KOWALCZYK, threads=8,
                                    want_fields=True)
    raw = np.float64 if prec == N.F64 else np.float32
    with Solver() as s:
        s.set_time_levels(2, 0)
        s.set_update_type(N.SRL)
        s.set_domain_labels(sc.labels)
        s.set_materials(sc.material_table())
        s.set_sources(sc.src_xyz, sc.src_kind, signal_f64=sc.signals,
                      signal_f32=sc.signals.astype(np.float32))
        s.set_receivers(sc.rec_xyz)
        s.run(sc.n_steps, prec)
        got = s.responses(raw)
        P, Pp = s.fields_raw(raw)
    assert np.abs(ref).max() > 1e-3
    assert np.array_equal(got, ref), f"max diff {np.abs(got - ref).max()}"
    assert np.array_equal(P.view(np.uint64 if raw == np.float64 else np.uint32),
                          P_ref.view(np.uint64 if raw == np.float64 else np.uint32))
    assert np.array_equal(Pp.view(np.uint64 if raw == np.float64 else np.uint32),
                          Pp_ref.view(np.uint64 if raw == np.float64 else np.uint32))


@pytest.mark.parametrize("levels", [1, 2])
@pytest.mark.parametrize("prec", [N.F32, N.F64, N.F16])
def test_centred_air_slices_bit_exact(gpu, prec, levels):
    """A room mostly of plain air (one box), so most wave tiles of most slices take the centred
    air path (centred_air: no byte decode, no table gather); fields compared bit for bit,
    signed zeros included, and responses against the oracle."""
    X, Y, Z = 190, 60, 24
    lab = rooms.box_room(X, Y, Z, floor=0, ceiling=1, walls=2)
    lab[6:14, 20:31, 70:95] = 3
    lab = fill_centred(lab)
    pos, mat, _ = oracle.classify_kow(lab)
    # most (slice, tile) windows of the two-level sweep are all centred air (0x80)
    air = pos == 0x80
    assert air[2:Z - 2, 2:Y - 2, 2:X - 2].mean() > 0.8
    rng = np.random.default_rng(7)
    n_steps = 48
    src = np.array([[40, 30, 12], [150, 12, 5], [120, 45, 18]], np.int32)
    kind = np.array([1, 0, 1], np.int32)
    sig = rng.uniform(-1, 1, (3, n_steps))
    sig[:, 20:] = 0.0
    rec = np.array([[45, 33, 12], [100, 20, 9], [180, 50, 20]], np.int32)
    table = np.zeros((4, 20), np.float32)
    table[:, 0] = [0.05, 0.2, 0.4, 0.7]
    pname, storage, scale, _ = PREC[prec]
    sg = sig if pname == "f64" else sig.astype(np.float32)
    ref, P_ref, Pp_ref = oracle.run(pos, mat, table, src, kind, sg, rec, n_steps, precision=pname,
                                    storage=storage, scale_exp=scale,
                                    literal=oracle.SCHEME_KOWALCZYK, threads=8, want_fields=True)
    raw = {N.F32: np.float32, N.F64: np.float64, N.F16: np.float16}[prec]
    bits = {N.F32: np.uint32, N.F64: np.uint64, N.F16: np.uint16}[prec]
    with Solver() as s:
        s.set_time_levels(levels, 0)
        s.set_update_type(N.SRL)
        s.set_domain_labels(lab)
        s.set_materials(table)
        s.set_sources(src, kind, signal_f64=sig, signal_f32=sig.astype(np.float32))
        s.set_receivers(rec)
        if scale:
            s.prepare_fields(prec, scale)
        s.run(n_steps, prec)
        got = s.responses(np.float64 if prec == N.F64 else np.float32)
        P, Pp = s.fields_raw(raw)
    assert np.abs(ref).max() > 1e-3
    assert np.array_equal(got, ref), f"max diff {np.abs(got - ref).max()}"
    if prec == N.F16:
        P_ref, Pp_ref = P_ref.astype(np.float16), Pp_ref.astype(np.float16)
    assert np.array_equal(P.view(bits), np.asarray(P_ref, raw).view(bits))
    assert np.array_equal(Pp.view(bits), np.asarray(Pp_ref, raw).view(bits))
