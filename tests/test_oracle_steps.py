"""Pins of the oracle's single-step mode (orc_step_calls_f): launchFDTD3dStep,
src/kernels/kernels3d.cu:376-482, called as App::executeStep calls it (App.cpp:407-436),
including the step-direction rule (:386, :462-465): the field pointers flip only while the
direction is unchanged, so the call that reverses time leaves the current field in place.

* all-forward calls are the reference's run loop (orc_run_f) bit for bit;
* forward K calls, then one backward call (step index K, no source sample there): the current
  field is still p^K and the past buffer holds p^{K+1}, bit for bit against forward runs of K
  and K+1 steps -- a rule that always flipped would leave p^{K+1} current;
* with rigid walls the leapfrog is time-reversible, p^{n-1} = L(p^n) - p^{n+1}: backward calls
  retrace the forward fields p^{K-1}, p^{K-2}, ... to rounding, and the response sample written
  at step index j is p^j at the receiver, i.e. the forward response sample j-1;
* a second reversal (backward, then forward again) keeps the pointers once more.
"""
import numpy as np
import pytest

import oracle
from parallelfdtd_amd import rooms


def scene(seed=3, dims=(15, 12, 11), n_steps=40, active=6):
    sc = rooms.scene_random_small(seed, dims=dims, n_mat=3, n_boxes=2, n_src=2, n_rec=3,
                                  n_steps=n_steps)
    pos, mat, _ = oracle.classify(sc.labels)
    sig = np.asarray(sc.signals, np.float32).copy()
    sig[:, active:] = 0.0  # sources silent after the first `active` steps
    sc.src_kind = np.ones(len(sc.src_kind), np.int32)  # soft: a silent step leaves the field
    return sc, pos, mat, sig


def fwd(sc, pos, mat, sig, table, n):
    return oracle.run(pos, mat, table, sc.src_xyz, sc.src_kind, sig, sc.rec_xyz, n,
                      want_fields=True)


def test_forward_calls_equal_the_run_loop():
    sc, pos, mat, sig = scene()
    n = 25
    table = sc.material_table()
    ref, P, Pp = fwd(sc, pos, mat, sig, table, n)
    got, Q, Qp = oracle.step_calls(pos, mat, table, sc.src_xyz, sc.src_kind, sig, sc.rec_xyz, n,
                                   np.arange(n), np.ones(n, np.int32))
    assert np.array_equal(got, ref) and np.array_equal(Q, P) and np.array_equal(Qp, Pp)


@pytest.mark.parametrize("K", [9, 20])
def test_reversing_call_keeps_the_current_field(K):
    sc, pos, mat, sig = scene()
    table = sc.material_table()
    n_steps = 40
    steps = list(range(K)) + [K]
    dirs = [1] * K + [-1]
    resp, Q, Qp = oracle.step_calls(pos, mat, table, sc.src_xyz, sc.src_kind, sig, sc.rec_xyz,
                                    n_steps, steps, dirs)
    rK, PK, _ = fwd(sc, pos, mat, sig, table, K)
    _, PK1, _ = fwd(sc, pos, mat, sig, table, K + 1)
    assert np.abs(PK).max() > 1e-3
    assert np.array_equal(Q, PK)    # no flip on the reversing call
    assert np.array_equal(Qp, PK1)  # the update it computed is the forward step K+1
    assert not np.array_equal(Q, PK1)
    # response sample K is the current field p^K: the forward run's sample K-1
    assert np.array_equal(resp[:, K], rK[:, K - 1])
    assert np.array_equal(resp[:, :K], rK[:, :K])


def test_backward_calls_retrace_with_rigid_walls():
    sc, pos, mat, sig = scene(seed=4)
    table = np.zeros(3, np.float32)  # rigid walls: the update is time-reversible
    K, back = 24, 14
    n_steps = 40
    steps = list(range(K)) + list(range(K, K - back, -1))
    dirs = [1] * K + [-1] * back
    resp, Q, Qp = oracle.step_calls(pos, mat, table, sc.src_xyz, sc.src_kind, sig, sc.rec_xyz,
                                    n_steps, steps, dirs)
    # after the reversing call (index K) and back-1 more: current = p^{K-back+1}
    j = K - back + 1
    fwd_resp, Pj, Pjm = fwd(sc, pos, mat, sig, table, j)
    _, Pj1, _ = fwd(sc, pos, mat, sig, table, j + 1)
    scale = np.abs(Pj).max()
    assert scale > 1e-3
    np.testing.assert_allclose(Q, Pj, rtol=0, atol=2e-5 * scale)
    np.testing.assert_allclose(Qp, Pj1, rtol=0, atol=2e-5 * scale)
    full, _, _ = fwd(sc, pos, mat, sig, table, K)
    for s in range(K - back + 1, K + 1):  # samples written while stepping back: p^s
        np.testing.assert_allclose(resp[:, s], full[:, s - 1], rtol=0, atol=2e-5 * scale)


def test_second_reversal_keeps_the_pointers_again():
    """forward K, backward 1 (index K), forward 1 (index K-1): the forward call after the
    backward one changes direction too, so it does not flip either."""
    sc, pos, mat, sig = scene(seed=6)
    table = sc.material_table()
    K = 15
    steps = list(range(K)) + [K, K - 1]
    dirs = [1] * K + [-1, 1]
    _, Q, Qp = oracle.step_calls(pos, mat, table, sc.src_xyz, sc.src_kind, sig, sc.rec_xyz, 40,
                                 steps, dirs)
    _, PK, _ = fwd(sc, pos, mat, sig, table, K)
    _, _, Pp2 = oracle.step_calls(pos, mat, table, sc.src_xyz, sc.src_kind, sig, sc.rec_xyz, 40,
                                  steps[:-1], dirs[:-1])
    # the third call updates (current p^K, past p^{K+1}) into the past buffer, no flip
    assert np.array_equal(Q, PK)
    assert not np.array_equal(Qp, Pp2)


def test_step_index_and_direction_checked():
    sc, pos, mat, sig = scene()
    table = sc.material_table()
    with pytest.raises(ValueError):
        oracle.step_calls(pos, mat, table, sc.src_xyz, sc.src_kind, sig, sc.rec_xyz, 10, [10], [1])
    with pytest.raises(ValueError):
        oracle.step_calls(pos, mat, table, sc.src_xyz, sc.src_kind, sig, sc.rec_xyz, 10, [0], [0])
