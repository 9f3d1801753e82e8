"""pfdtd_capture_slice_noalloc (the capture entry point that allocates no device memory) against
pfdtd_capture_slice through the C ABI: two calls in a row on one GPU, the second with a smaller
plane, in FP32 and FP16 storage; planes above PFDTD_STATIC_CAPTURE_MAX are refused."""
import numpy as np
import pytest

import oracle
from parallelfdtd_amd import native as N
from parallelfdtd_amd import rooms
from parallelfdtd_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _captures(sc, prec, scale, planes):
    """(allocating, static) captures of each (slice, orientation) after the scene's run."""
    out = []
    with Solver() as s:
        s.load_scene(sc)
        if scale:
            s.prepare_fields(prec, scale)
        s.run(sc.n_steps, prec)
        for sl, orient in planes:
            out.append((s.capture_slice(sl, orient), s.capture_slice_noalloc(sl, orient)))
    return out


@pytest.mark.parametrize("prec,scale", [(N.F32, 0), (N.F16, 8)])
def test_noalloc_capture_equals_capture(gpu, prec, scale):
    big = rooms.scene_random_small(81, dims=(300, 140, 40), n_mat=3, n_boxes=6, n_src=3,
                                   n_rec=2, n_steps=70)
    small = rooms.scene_random_small(82, dims=(37, 21, 15), n_mat=2, n_boxes=2, n_src=2,
                                     n_rec=2, n_steps=30)
    first = _captures(big, prec, scale, [(20, 0), (70, 1)])
    second = _captures(small, prec, scale, [(7, 0), (3, 2)])
    for (a, b) in first + second:
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert np.abs(a[0]).max() > 0
    # the second call's smaller plane: nothing of the first call's larger plane leaks into it
    assert second[0][1][0].shape == (21, 37)


def test_noalloc_capture_matches_oracle_and_refuses(gpu):
    sc = rooms.scene_random_small(83, dims=(45, 31, 19), n_steps=40)
    pos, mat, _ = oracle.classify(sc.labels)
    _, P_ref, _ = oracle.run(pos, mat, sc.material_table(), sc.src_xyz, sc.src_kind,
                             sc.signals.astype(np.float32), sc.rec_xyz, sc.n_steps,
                             want_fields=True)
    with Solver() as s:
        s.load_scene(sc)
        s.run(sc.n_steps, N.F32)
        for orient, lim in ((0, 19), (1, 31), (2, 45)):
            for sl in (0, lim // 2, lim - 1):
                p, q = s.capture_slice_noalloc(sl, orient)
                ep, eq = oracle.capture_slice(P_ref, pos, sl, orient)
                assert np.array_equal(p, ep) and np.array_equal(q, eq)
            with pytest.raises(N.PfdtdError) as e:
                s.capture_slice_noalloc(lim, orient)
            assert e.value.code == N.E_ARG
    # a plane one row above the static buffer's size is refused before anything runs
    X = 2050
    Y = N.STATIC_CAPTURE_MAX // (X - 0) + 1
    lab = rooms.box_room(X, Y, 3)
    assert X * Y > N.STATIC_CAPTURE_MAX
    with Solver() as s:
        s.set_domain_labels(lab)
        s.set_materials(np.array([[0.1]], np.float32))
        s.prepare_fields(N.F32, 0)
        with pytest.raises(N.PfdtdError) as e:
            s.capture_slice_noalloc(1, 0)
        assert e.value.code == N.E_ARG
        p, _ = s.capture_slice_noalloc(1, 1)  # the xz plane of the same room fits
        assert p.shape == (3, X) and not p.any()
