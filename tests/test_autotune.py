"""The autotuner's bookkeeping on the CPU: a stand-in solver returns fixed step times, and the
sweep must time single-level candidates for every batch width in the interleaved layout,
two-level candidates (both block shapes, every z-chunk) and three-level candidates (every
z-chunk) unbatched for B = 1 and in the grouped layout for B > 1, restore the single-level interleaved setting between them, report the best
per B with and without the two-level path, and choose the fastest per source."""
import numpy as np
import pytest

from parallelfdtd_amd import autotune
from parallelfdtd_amd import native as N
from parallelfdtd_amd import rooms


class FakeSolver:
    """Records the settings of each timed run; the step time is a function of them."""

    def __init__(self, device=0, times=None):
        self.levels, self.zc, self.rows, self.tz, self.shape, self.B = 1, 0, 2, 8, N.SWEEP_AUTO, 1
        self.layout = N.BATCH_INTERLEAVED
        self.times = times
        self.runs = []

    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False

    def set_domain_labels(self, labels):
        pass

    def set_materials(self, table):
        pass

    def set_sources(self, xyz, kind, signal_f64=None, signal_f32=None):
        pass

    def set_receivers(self, xyz):
        pass

    def set_batch(self, B):
        self.B = B

    def set_batch_layout(self, layout):
        self.layout = layout

    def set_source_slots(self, s):
        pass

    def set_receiver_slots(self, s):
        pass

    def set_time_levels(self, levels=0, zchunk=0):
        self.levels, self.zc = levels, zchunk

    def set_sweep_shape(self, shape=N.SWEEP_AUTO):
        self.shape = shape

    def set_update_tiling(self, rows, zchunk):
        self.rows, self.tz = rows, zchunk

    def run(self, n_steps, precision):
        key = (self.B, self.layout, self.levels, self.rows if self.levels == 1 else self.shape,
               self.tz if self.levels == 1 else self.zc)
        self.runs.append(key)
        return n_steps, self.times(key) * n_steps


def sweep_with(times, batches=(1, 8, 32)):
    scene = rooms.scene_c4(n_sources=max(batches), n_steps=4, scale=16)
    holder = {}

    def make(device=0):
        holder["s"] = FakeSolver(device, times)
        return holder["s"]

    orig = autotune.Solver
    autotune.Solver = make
    try:
        rep = autotune.sweep(scene, batches=batches, rows=(1, 2), zchunks=(4, 8), steps=4,
                             warmup=1, log=lambda m: None)
    finally:
        autotune.Solver = orig
    return rep, holder["s"], scene.n_nodes


def test_sweep_candidates_and_batched_runs_single_level():
    rep, s, _ = sweep_with(lambda key: 1.0 + key[0])
    two = [c for c in rep["candidates"] if c["levels"] == 2]
    assert {c["B"] for c in two} == {1, 8, 32}
    grid = sorted((name, zc) for name in autotune.SHAPES for zc in autotune.TWO_LEVEL_ZCHUNKS)
    for B in (1, 8, 32):
        mine = [c for c in two if c["B"] == B]
        assert sorted((c["shape"], c["zchunk"]) for c in mine) == grid
        assert {c["layout"] for c in mine} == {"interleaved" if B == 1 else "grouped"}
    three = [c for c in rep["candidates"] if c["levels"] == 3]
    for B in (1, 8, 32):
        mine = [c for c in three if c["B"] == B]
        assert sorted(c["zchunk"] for c in mine) == list(autotune.THREE_LEVEL_ZCHUNKS)
        assert {c["layout"] for c in mine} == {"interleaved" if B == 1 else "grouped"}
    # interleaved batched runs (B > 1) are timed with one time level, grouped ones with two or
    # three
    assert all(levels == 1 for (B, lay, levels, _, _) in s.runs
               if B > 1 and lay == N.BATCH_INTERLEAVED)
    assert all(levels in (2, 3) for (B, lay, levels, _, _) in s.runs if lay == N.BATCH_GROUPED)
    assert len([c for c in rep["candidates"] if c["levels"] == 1]) == 3 * 2 * 2


def test_choice_prefers_two_level_when_faster():
    def t(key):
        B, lay, levels, tile, zc = key
        if levels == 2:
            per = 0.5 if (tile, zc) == (N.SWEEP_COLUMN, 8) else 0.9
            return per if B == 1 else 0.55 * B  # grouped: slightly slower per source
        return 1.0 * B  # batching gives no per-source gain here
    rep, _, nodes = sweep_with(t)
    assert rep["choice"] == {"B": 1, "layout": "interleaved", "levels": 2, "rows": None,
                             "zchunk": 8, "shape": "column"}
    assert rep["best"]["1"]["levels"] == 2
    assert rep["best_single_level"]["1"]["levels"] == 1
    assert rep["best"]["32"]["layout"] == "grouped"
    assert rep["best"]["32"]["per_source_speedup_vs_B1"] == pytest.approx(0.5 / 0.55, rel=1e-3)
    assert rep["best"]["1"]["ms_per_step"] == pytest.approx(0.5)
    assert rep["best"]["1"]["per_source_gvox"] == pytest.approx(nodes / 0.5e-3 / 1e9, rel=1e-5)


def test_choice_prefers_batching_when_faster():
    def t(key):
        B, lay, levels, tile, zc = key
        base = 1.0 if levels == 1 else 0.8
        return base * B / (3.0 if B == 32 and tile == 2 and zc == 4 else 1.0)
    rep, _, _ = sweep_with(t)
    assert rep["choice"] == {"B": 32, "layout": "interleaved", "levels": 1, "rows": 2,
                             "zchunk": 4, "shape": None}
    assert rep["best_single_level"]["32"]["per_source_speedup_vs_B1"] == pytest.approx(3.0, rel=1e-3)
    assert rep["best"]["32"]["per_source_speedup_vs_B1"] == pytest.approx(2.4, rel=1e-3)


def test_choice_prefers_grouped_two_level_when_faster():
    def t(key):
        B, lay, levels, tile, zc = key
        if lay == N.BATCH_GROUPED:
            return 0.6 * B if (tile, zc) == (N.SWEEP_WIDE, 16) else 0.9 * B
        return (0.8 if levels == 2 else 1.0) * B
    rep, _, _ = sweep_with(t)
    assert rep["choice"] == {"B": 8, "layout": "grouped", "levels": 2, "rows": None,
                             "zchunk": 16, "shape": "wide"}
    assert rep["best"]["32"]["per_source_speedup_vs_B1"] == pytest.approx(0.8 / 0.6, rel=1e-3)
