"""GPU parity of the two-time-level update (temporal blocking) against the CPU oracle.

The oracle advances one step at a time; the GPU advances pairs of steps per sweep with source
injection and receiver gather inside the kernel. Bar: bit-identical responses and final fields
in every precision, for odd and even step counts and several z-chunk lengths."""
import numpy as np
import pytest

import oracle
from parallelfdtd_amd import native as N
from parallelfdtd_amd import rooms
from parallelfdtd_amd.solver import Solver

pytestmark = pytest.mark.gpu

PREC = {N.F32: ("f32", 0, 0, np.float32), N.F64: ("f64", 0, 0, np.float64),
        N.F16: ("f32", oracle.STORE_F16, 9, np.float16), N.BF16: ("f32", oracle.STORE_BF16, 9, np.uint16)}


def run_both(sc, prec, levels=2, zchunk=32, literal=0, octave=0, poll=100, shape=N.SWEEP_AUTO):
    pname, storage, scale, raw = PREC[prec]
    pos, mat, _ = oracle.classify(sc.labels)
    sig = sc.signals if pname == "f64" else sc.signals.astype(np.float32)
    ref, P_ref, Pp_ref = oracle.run(pos, mat, sc.material_table(), sc.src_xyz, sc.src_kind, sig,
                                    sc.rec_xyz, sc.n_steps, precision=pname, octave=octave,
                                    literal=literal, storage=storage, scale_exp=scale,
                                    want_fields=True)
    with Solver() as s:
        s.set_time_levels(levels, zchunk)
        s.set_sweep_shape(shape)
        s.set_domain_labels(sc.labels)
        s.set_materials(sc.material_table(), octave=octave, literal_slot=bool(literal & 1),
                        soft_overwrite=bool(literal & 2))
        s.set_sources(sc.src_xyz, sc.src_kind, signal_f64=sc.signals,
                      signal_f32=sc.signals.astype(np.float32))
        s.set_receivers(sc.rec_xyz)
        s.set_poll_interval(poll)
        if scale:
            s.prepare_fields(prec, scale)
        done, _ = s.run(sc.n_steps, prec)
        got = s.responses(np.float64 if pname == "f64" else np.float32)
        P, Pp = s.fields_raw(raw)
    assert done == sc.n_steps
    if storage:  # raw half-precision fields: compare through the oracle's stored values
        conv = (lambda h: h.astype(np.float32)) if raw == np.float16 else \
               (lambda h: (h.astype(np.uint32) << 16).view(np.float32))
        P, Pp = conv(P), conv(Pp)  # both sides hold fields scaled by 2^scale
    return got, ref, (P, Pp), (P_ref, Pp_ref)


def assert_same(got, ref, F, F_ref):
    assert np.array_equal(got, ref), f"responses differ: max {np.abs(got - ref).max()}"
    assert np.array_equal(F[0], F_ref[0]) and np.array_equal(F[1], F_ref[1]), "fields differ"


@pytest.mark.parametrize("prec", [N.F32, N.F64, N.F16, N.BF16])
@pytest.mark.parametrize("n_steps,zchunk", [(64, 32), (61, 5), (40, 1)])
def test_two_level_bit_exact(gpu, prec, n_steps, zchunk):
    sc = rooms.scene_random_small(81, dims=(130, 70, 29), n_mat=4, n_boxes=6, n_src=3, n_rec=6,
                                  n_steps=n_steps)
    sc.rec_xyz[0] = sc.src_xyz[0]  # a receiver on a source voxel reads the pre-injection value
    assert_same(*run_both(sc, prec, zchunk=zchunk))


@pytest.mark.parametrize("literal", [1, 2, 3])
def test_two_level_quirk_modes(gpu, literal):
    sc = rooms.scene_random_small(82, dims=(70, 40, 20), n_mat=3, n_boxes=4, n_src=4, n_rec=4,
                                  n_steps=50)
    assert_same(*run_both(sc, N.F32, literal=literal, octave=1 if literal & 1 else 0))


@pytest.mark.parametrize("levels", [2, 3])
def test_two_level_crowded_sources_and_receivers(gpu, levels):
    """More sources/receivers in one wave tile than its list holds (scan-all path), several
    sources on one voxel (applied in index order), hard and soft mixed; two and three levels."""
    sc = rooms.scene_random_small(83, dims=(66, 36, 16), n_mat=2, n_boxes=0, n_src=1, n_rec=1,
                                  n_steps=30)
    pts = np.array([[10 + i % 6, 5 + (i // 6) % 5, 4 + i // 30] for i in range(45)], np.int32)
    src = np.concatenate([pts, pts[:3]])
    kinds = np.array([i % 2 for i in range(len(src))], np.int32)
    rng = np.random.default_rng(3)
    sc.src_xyz, sc.src_kind = src, kinds
    sc.signals = rng.uniform(-1, 1, (len(src), sc.n_steps))
    sc.rec_xyz = pts[::-1].copy()
    for prec in (N.F32, N.F64):
        assert_same(*run_both(sc, prec, levels=levels, zchunk=4))


@pytest.mark.parametrize("levels", [2, 3])
def test_two_level_falls_back_for_boundary_slice_source(gpu, levels):
    sc = rooms.scene_random_small(84, dims=(40, 30, 20), n_steps=30)
    sc.src_xyz[0] = [5, 5, 0]  # on the never-updated slice z = 0
    assert_same(*run_both(sc, N.F32, levels=levels))


def test_two_level_interrupt(gpu):
    sc = rooms.scene_random_small(85, dims=(40, 30, 20), n_steps=60)
    calls = []
    with Solver() as s:
        s.load_scene(sc)
        s.set_poll_interval(10)
        s.run(sc.n_steps, N.F32)
        full = s.responses(np.float32)
        done, _ = s.run(sc.n_steps, N.F32, interrupt=lambda: len(calls) >= 3 or calls.append(1))
        part = s.responses(np.float32)
    assert done == 30
    assert np.array_equal(part[:, :30], full[:, :30]) and not part[:, 30:].any()


def test_levels_agree_on_c1(gpu):
    """Config 1 through single steps, two and three levels per sweep: identical bits."""
    sc = rooms.scene_c1(n_steps=600)
    out = []
    for levels in (1, 2, 3):
        with Solver() as s:
            s.set_time_levels(levels)
            s.load_scene(sc)
            s.run(sc.n_steps, N.F32)
            out.append(s.responses(np.float32))
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])


@pytest.mark.parametrize("prec", [N.F32, N.F16])
def test_two_level_sources_on_tile_edges(gpu, prec):
    """Sources and receivers on the columns and rows where 64 x 16 tiles meet (their one-voxel
    rings overlap there), including the last partial tile of each axis."""
    sc = rooms.scene_random_small(86, dims=(131, 69, 23), n_mat=3, n_boxes=3, n_src=1, n_rec=1,
                                  n_steps=41)
    xs = [1, 62, 63, 64, 65, 127, 128, 129]
    ys = [1, 15, 16, 17, 31, 32, 47, 48, 64, 67]
    pts = np.array([[xs[i % len(xs)], ys[(3 * i) % len(ys)], 2 + (5 * i) % 19] for i in range(40)],
                   np.int32)
    rng = np.random.default_rng(7)
    sc.src_xyz = pts[:24].copy()
    sc.src_kind = np.array([i % 2 for i in range(24)], np.int32)
    sc.signals = rng.uniform(-1, 1, (24, sc.n_steps))
    sc.rec_xyz = pts[::-1].copy()
    assert_same(*run_both(sc, prec, zchunk=7))


def border_voxel_scene(seed, dims=(185, 80, 26), n_steps=40):
    """Single solid voxels two and one columns/rows either side of every wave-tile border (x
    tiles of 60 columns; y tiles of 4, 6 and 8 rows): each puts boundary voxels on exactly one
    side of the border, so a tile whose air-row bits left out its halo column or row would
    update a boundary voxel as air."""
    sc = rooms.scene_random_small(seed, dims=dims, n_mat=3, n_boxes=0, n_src=1, n_rec=1,
                                  n_steps=n_steps)
    X, Y, Z = dims
    xs = sorted({60 * t + d for t in range(1, (X + 59) // 60) for d in (-2, -1, 0, 1)} & set(range(2, X - 2)))
    ys = sorted({m + d for ty in (4, 6, 8) for m in range(ty, Y, ty) for d in (-2, -1, 0, 1)}
                & set(range(2, Y - 2)))
    k = 0
    for x in xs:
        for y in ys:
            z = 2 + (k * 7) % (Z - 4)
            sc.labels[z, y, x] = 2 + k % 2
            k += 1
    rng = np.random.default_rng(seed)
    n_src = 48
    sc.src_xyz = np.stack([rng.integers(2, X - 2, n_src), rng.integers(2, Y - 2, n_src),
                           rng.integers(2, Z - 2, n_src)], axis=1).astype(np.int32)
    sc.src_kind = np.ones(n_src, np.int32)
    sc.signals = rng.uniform(-1, 1, (n_src, n_steps))
    sc.rec_xyz = sc.src_xyz[::3].copy()
    return sc


@pytest.mark.parametrize("prec", [N.F32, N.F64, N.F16, N.BF16])
def test_two_level_boundary_voxels_on_tile_borders(gpu, prec):
    sc = border_voxel_scene(87)
    got, ref, F, F_ref = run_both(sc, prec, zchunk=9)
    assert np.abs(F_ref[0]).max() > 0
    assert_same(got, ref, F, F_ref)


@pytest.mark.parametrize("prec", [N.F32, N.F64])
def test_two_level_wide_room_block_shapes(gpu, prec):
    """A room 961 columns wide (17 wave tiles of 60): FP32 takes blocks of 4 x 4 wave tiles, FP64
    blocks of 2 x 2, each with tiles past the grid in the last block column; boundary voxels on
    the tile borders, bit-exact against the oracle."""
    sc = border_voxel_scene(88, dims=(961, 34, 12), n_steps=24)
    got, ref, F, F_ref = run_both(sc, prec, zchunk=0)
    assert np.abs(F_ref[0]).max() > 0
    assert_same(got, ref, F, F_ref)


def test_two_level_domain_reset_with_same_voxel_count(gpu):
    """One context, three domains: an empty room, the same shape with border voxels, and the
    transposed shape (same voxel count): each run equals the oracle, so the packed plane and
    its air-row bits follow the domain."""
    a = rooms.scene_random_small(88, dims=(122, 60, 20), n_mat=3, n_boxes=0, n_src=4, n_rec=4,
                                 n_steps=24)
    b = border_voxel_scene(89, dims=(122, 60, 20), n_steps=24)
    c = border_voxel_scene(90, dims=(60, 122, 20), n_steps=24)
    with Solver() as s:
        s.set_time_levels(2, 8)
        for sc in (a, b, c):
            pos, mat, _ = oracle.classify(sc.labels)
            sig = sc.signals.astype(np.float32)
            ref = oracle.run(pos, mat, sc.material_table(), sc.src_xyz, sc.src_kind, sig,
                             sc.rec_xyz, sc.n_steps)
            s.set_domain_labels(sc.labels)
            s.set_materials(sc.material_table())
            s.set_sources(sc.src_xyz, sc.src_kind, signal_f64=sc.signals, signal_f32=sig)
            s.set_receivers(sc.rec_xyz)
            s.run(sc.n_steps, N.F32)
            assert np.array_equal(s.responses(np.float32), ref)


@pytest.mark.parametrize("shape", [N.SWEEP_COLUMN, N.SWEEP_WIDE])
@pytest.mark.parametrize("prec", [N.F32, N.F64, N.F16])
def test_two_level_block_shapes(gpu, shape, prec):
    """Both block shapes on a room whose 4 x tiles leave wide blocks partly past the grid (FP64
    and the 16-bit sweep always use columns), with boundary voxels on the tile borders."""
    sc = border_voxel_scene(91, dims=(250, 54, 20), n_steps=30)
    assert_same(*run_both(sc, prec, zchunk=7, shape=shape))


def test_two_level_wide_room_automatic_shape(gpu):
    """A room of 1000 columns (17 x tiles) takes wide blocks by default; 3 of its 20 block tiles
    lie past the last column."""
    sc = rooms.scene_random_small(92, dims=(1000, 30, 12), n_mat=3, n_boxes=5, n_src=6, n_rec=8,
                                  n_steps=26)
    assert_same(*run_both(sc, N.F32, zchunk=4))
    assert_same(*run_both(sc, N.F32, zchunk=4, shape=N.SWEEP_COLUMN))


def test_sweep_shape_rejects_unknown(gpu):
    with Solver() as s:
        with pytest.raises(N.PfdtdError):
            s.set_sweep_shape(4)
        with pytest.raises(N.PfdtdError):
            s.set_sweep_shape(-1)


def pair_border_scene(seed, dims, n_steps=30):
    """Single solid voxels next to the 124-column borders of the two-column sweep's wave tiles
    (and row borders of 4-, 6- and 7-row tiles), and sources and receivers on both columns of the pairs either
    side of those borders (lanes 0, 1, 62 and 63 of a wave) and on the room's last columns."""
    sc = rooms.scene_random_small(seed, dims=dims, n_mat=3, n_boxes=0, n_src=1, n_rec=1,
                                  n_steps=n_steps)
    X, Y, Z = dims
    borders = {124 * t + d for t in range(1, X // 124 + 2) for d in (-3, -2, -1, 0, 1, 2)}
    xs = sorted((borders | {2, 3, X - 4, X - 3}) & set(range(2, X - 2)))
    ys = sorted({m + d for ty in (4, 6, 7) for m in range(ty, Y, ty) for d in (-2, -1, 0, 1)}
                & set(range(2, Y - 2)))
    k = 0
    for x in xs:
        for y in ys[::2]:
            z = 2 + (k * 5) % max(1, Z - 4)
            sc.labels[z, y, x] = 2 + k % 2
            k += 1
    pts = np.array([[xs[i % len(xs)], ys[(7 * i) % len(ys)], 1 + (3 * i) % (Z - 2)]
                    for i in range(2 * len(xs))], np.int32)
    rng = np.random.default_rng(seed)
    n_src = min(len(pts), 40)
    sc.src_xyz = pts[:n_src].copy()
    sc.src_kind = np.array([i % 2 for i in range(n_src)], np.int32)
    sc.signals = rng.uniform(-1, 1, (n_src, n_steps))
    sc.rec_xyz = pts[::-1].copy()
    return sc


@pytest.mark.parametrize("dims", [(250, 40, 18), (372, 27, 14), (126, 25, 12), (8, 9, 7),
                                  (131, 22, 10), (500, 45, 9)])
@pytest.mark.parametrize("shape", [N.SWEEP_AUTO, N.SWEEP_PAIR])
@pytest.mark.parametrize("prec", [N.F16, N.BF16])
def test_two_level_pair_sweep(gpu, dims, shape, prec):
    """The two-column 16-bit sweep (automatic on rooms of even X) on rooms of two tiles and a
    pair, exactly three tiles, one tile and a pair, less than one tile, odd X (where PAIR falls
    back to one column per lane) and five tiles (three block columns, the last holding one x
    tile over the rows of two blocks): bit-exact against the oracle, boundary voxels, sources
    and receivers on the pair and tile borders."""
    sc = pair_border_scene(95, dims)
    got, ref, F, F_ref = run_both(sc, prec, zchunk=5, shape=shape)
    assert np.abs(ref).max() > 0
    assert_same(got, ref, F, F_ref)


@pytest.mark.parametrize("prec", [N.F32, N.F16])
def test_two_level_automatic_zchunk(gpu, prec):
    """z-chunk 0 (the default) picks the chunk length from the grid; results do not change."""
    sc = rooms.scene_random_small(93, dims=(130, 70, 41), n_mat=4, n_boxes=6, n_src=3, n_rec=6,
                                  n_steps=33)
    assert_same(*run_both(sc, prec, zchunk=0))
    with Solver() as s:
        with pytest.raises(N.PfdtdError):
            s.set_time_levels(2, -1)
        with pytest.raises(N.PfdtdError):
            s.set_time_levels(4, 0)
        s.set_time_levels(3, 0)
        s.set_update_tiling(0, 0)  # back to the automatic single-step tiling
        with pytest.raises(N.PfdtdError):
            s.set_update_tiling(0, 8)
        with pytest.raises(N.PfdtdError):
            s.set_update_tiling(3, 8)


@pytest.mark.parametrize("dims", [(3, 3, 3), (4, 5, 3), (61, 3, 4), (3, 70, 5), (121, 33, 6),
                                  (241, 7, 9)])
@pytest.mark.parametrize("prec", [N.F32, N.F16])
def test_two_level_degenerate_shapes(gpu, dims, prec):
    """Rooms down to one interior voxel, widths of one tile plus a column, heights below one
    wave tile, and (241 columns) a wide block with three of its four tiles past the grid."""
    X, Y, Z = dims
    sc = rooms.scene_random_small(94, dims=dims, n_mat=2, n_boxes=0, n_src=1, n_rec=1,
                                  n_steps=13)
    sc.src_xyz = np.array([[X // 2, Y // 2, Z // 2]], np.int32)
    sc.src_kind = np.array([0], np.int32)
    sc.rec_xyz = np.array([[X // 2, Y // 2, Z // 2], [1, 1, 1], [X - 2, Y - 2, Z - 2]], np.int32)
    for shape in (N.SWEEP_COLUMN, N.SWEEP_WIDE):
        got, ref, F, F_ref = run_both(sc, prec, zchunk=0, shape=shape)
        assert np.abs(ref).max() > 0
        assert_same(got, ref, F, F_ref)


@pytest.mark.parametrize("lam", [0.5, 0.57735])
def test_two_level_pair_sweep_other_lambda(gpu, lam):
    """With lambda^2 not rounding to 1/3 the air entry is not (1, 1, 0): the two-column sweep
    then updates every slice through the coefficient table; bit-exact against the oracle."""
    sc = pair_border_scene(96, (250, 30, 12), n_steps=24)
    pos, mat, _ = oracle.classify(sc.labels)
    sig = sc.signals.astype(np.float32)
    ref, P_ref, _ = oracle.run(pos, mat, sc.material_table(), sc.src_xyz, sc.src_kind, sig,
                               sc.rec_xyz, sc.n_steps, lam=lam, storage=oracle.STORE_F16,
                               scale_exp=9, want_fields=True)
    with Solver() as s:
        s.set_domain_labels(sc.labels)
        s.set_materials(sc.material_table(), lam=lam)
        s.set_sources(sc.src_xyz, sc.src_kind, signal_f64=sc.signals, signal_f32=sig)
        s.set_receivers(sc.rec_xyz)
        s.prepare_fields(N.F16, 9)
        s.run(sc.n_steps, N.F16)
        got = s.responses(np.float32)
        P, _ = s.fields_raw(np.float16)
    assert np.abs(ref).max() > 0
    assert np.array_equal(got, ref)
    assert np.array_equal(P.astype(np.float32), P_ref)


def three_level_scene(seed, dims, n_steps):
    """Single solid voxels next to the 58-column and 4-row borders of the three-level sweep's
    wave tiles, sources and receivers on the columns either side of those borders (lanes 0-5
    and 58-63 of a wave) and near the room's last columns and rows."""
    sc = rooms.scene_random_small(seed, dims=dims, n_mat=3, n_boxes=2, n_src=1, n_rec=1,
                                  n_steps=n_steps)
    X, Y, Z = dims
    xs = sorted({58 * t + d for t in range(1, X // 58 + 2) for d in (-4, -3, -2, -1, 0, 1, 2, 3)}
                | {2, 3, 4, X - 5, X - 4, X - 3})
    xs = [x for x in xs if 2 <= x <= X - 3]
    ys = sorted({m + d for m in range(4, Y, 4) for d in (-2, -1, 0, 1)} & set(range(2, Y - 2)))
    k = 0
    for x in xs:
        for y in ys[::3]:
            sc.labels[2 + (k * 5) % max(1, Z - 4), y, x] = 2 + k % 2
            k += 1
    pts = np.array([[xs[i % len(xs)], ys[(5 * i) % len(ys)], 1 + (3 * i) % (Z - 2)]
                    for i in range(2 * len(xs))], np.int32)
    rng = np.random.default_rng(seed)
    n_src = min(len(pts), 30)
    sc.src_xyz = pts[:n_src].copy()
    sc.src_kind = np.array([i % 2 for i in range(n_src)], np.int32)
    sc.signals = rng.uniform(-1, 1, (n_src, n_steps))
    sc.rec_xyz = pts[::-1].copy()
    return sc


@pytest.mark.parametrize("dims,n_steps,zchunk", [((180, 41, 19), 31, 5), ((121, 23, 12), 30, 1),
                                                 ((59, 9, 8), 29, 0), ((7, 6, 5), 16, 2),
                                                 ((300, 30, 26), 24, 0)])
def test_three_level_bit_exact(gpu, dims, n_steps, zchunk):
    """Three time levels per sweep (FP32): bit-exact against the oracle and against two levels,
    on rooms of several 58-column tiles, a tile and a column, one tile and one column past it,
    below one tile, and with z-chunks of 1, 2 and 5 slices; step counts that leave 1, 0 and 2
    steps after the last three-level launch."""
    sc = three_level_scene(97, dims, n_steps)
    got, ref, F, F_ref = run_both(sc, N.F32, levels=3, zchunk=zchunk)
    assert np.abs(ref).max() > 0
    assert_same(got, ref, F, F_ref)
    got2, _, F2, _ = run_both(sc, N.F32, levels=2, zchunk=zchunk)
    assert np.array_equal(got.view(np.uint32), got2.view(np.uint32))
    assert np.array_equal(F[0].view(np.uint32), F2[0].view(np.uint32))


@pytest.mark.parametrize("dims", [(3, 3, 3), (4, 5, 3), (61, 3, 4), (3, 70, 5), (121, 33, 6),
                                  (59, 9, 4)])
@pytest.mark.parametrize("prec", [N.F32, N.F64])
def test_three_level_degenerate_shapes(gpu, dims, prec):
    """Three levels on rooms down to one interior voxel, one or two interior slices, heights
    below one wave tile and widths of one tile plus a column."""
    X, Y, Z = dims
    sc = rooms.scene_random_small(94, dims=dims, n_mat=2, n_boxes=0, n_src=1, n_rec=1,
                                  n_steps=13)
    sc.src_xyz = np.array([[X // 2, Y // 2, Z // 2]], np.int32)
    sc.src_kind = np.array([0], np.int32)
    sc.rec_xyz = np.array([[X // 2, Y // 2, Z // 2], [1, 1, 1], [X - 2, Y - 2, Z - 2]], np.int32)
    got, ref, F, F_ref = run_both(sc, prec, levels=3, zchunk=0)
    assert np.abs(ref).max() > 0
    assert_same(got, ref, F, F_ref)


def test_three_level_column_blocks(gpu):
    """The three-level sweep in blocks of 1 x 4 wave tiles (PFDTD_SWEEP_COLUMN)."""
    sc = three_level_scene(96, (180, 41, 19), 27)
    assert_same(*run_both(sc, N.F32, levels=3, zchunk=6, shape=N.SWEEP_COLUMN))


@pytest.mark.parametrize("literal", [1, 2])
def test_three_level_quirks_and_poll(gpu, literal):
    """The literal material slot and overwriting soft sources, with progress polls every 7
    steps (bundles that end between three-level launches)."""
    sc = three_level_scene(98, (130, 30, 14), 40)
    assert_same(*run_both(sc, N.F32, levels=3, zchunk=4, literal=literal,
                          octave=1 if literal & 1 else 0, poll=7))


@pytest.mark.parametrize("prec", [N.F64, N.F16, N.BF16])
def test_three_level_other_precisions(gpu, prec):
    """Three time levels with FP64 fields and with 16-bit storage (scaled, rounded at every
    level): bit-exact against the oracle."""
    sc = three_level_scene(99, (124, 20, 10), 22)
    assert_same(*run_both(sc, prec, levels=3, zchunk=3))
