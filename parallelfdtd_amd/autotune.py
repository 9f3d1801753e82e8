"""Autotuner for batched runs: sweeps the batch width B (sources advanced per launch, one slot
each) in both field layouts -- interleaved source-innermost, advanced by the single-level
batched update (its tile: y rows per thread, a 256-thread block covering 256/B x-columns of one
row band, and the z-chunk length), and grouped source-outermost, advanced by the two-level
sweep (block shape and z-chunk) and the three-level sweep (z-chunk) in one launch for all
slots -- timing a short run of each candidate on the real domain.

Throughput is per source: voxel-updates/s summed over the B slots,
  per_source_gvox = N_nodes * B * steps / seconds / 1e9,
with the step loop's own HIP-event time (injection and gather included). B = 1 is timed both
as the unbatched single-level update (the batched kernel's own baseline) and as the default
unbatched paths, the two-time-level sweep (z-chunks `TWO_LEVEL_ZCHUNKS`, column and wide blocks)
and the three-time-level sweep (z-chunks `THREE_LEVEL_ZCHUNKS`, blocks of 2 x 2 wave tiles); the
best of all candidates is the tuner's choice.

Run:  python -m parallelfdtd_amd.autotune --batches 1,8,16,32,64 --out tune_c4.json
"""
from __future__ import annotations

import argparse
import json
import sys
from pathlib import Path

import numpy as np

from . import native as N
from . import rooms
from .solver import Solver

BYTES_F32 = {"per_voxel": 2, "per_slot": 12}  # 2 + 12*B bytes per voxel-step, FP32 fields
TWO_LEVEL_ZCHUNKS = (8, 16, 32)
THREE_LEVEL_ZCHUNKS = (16, 32)
SHAPES = {"column": N.SWEEP_COLUMN, "wide": N.SWEEP_WIDE}


def load_batch(s: Solver, scene, B: int, rec_per_src: int) -> None:
    """Domain of `scene` with its first B sources, one per slot, and their receivers."""
    s.set_sources(scene.src_xyz[:B], scene.src_kind[:B], signal_f64=scene.signals[:B],
                  signal_f32=scene.signals[:B].astype(np.float32))
    s.set_receivers(scene.rec_xyz[: B * rec_per_src])
    s.set_batch(B)
    s.set_source_slots(np.arange(B, dtype=np.int32))
    s.set_receiver_slots(np.repeat(np.arange(B, dtype=np.int32), rec_per_src))


def time_config(s: Solver, steps: int, warmup: int, precision: int) -> float:
    """Milliseconds per step of the step loop (after a warm-up run)."""
    if warmup:
        s.run(warmup, precision)
    done, ms = s.run(steps, precision)
    if done != steps:
        raise RuntimeError(f"run stopped after {done} of {steps} steps")
    return ms / steps


def sweep(scene, batches=(1, 8, 16, 32, 64), rows=(1, 2, 4, 8), zchunks=(4, 8, 16),
          steps: int = 20, warmup: int = 3, rec_per_src: int = 8, device: int = 0,
          log=print) -> dict:
    nodes = scene.n_nodes
    if max(batches) > len(scene.src_xyz):
        raise ValueError(f"scene has {len(scene.src_xyz)} sources, batch {max(batches)} asked")
    out = {"scene": scene.name, "nodes": nodes, "steps": steps, "candidates": [], "best": {},
           "best_single_level": {}}

    def record(B, levels, ry, zc, ms, shape=None, layout="interleaved"):
        per_src = nodes * B / (ms * 1e-3) / 1e9
        gbs = nodes * (BYTES_F32["per_voxel"] + BYTES_F32["per_slot"] * B) / (ms * 1e-3) / 1e9
        # six significant digits: the speedups below are ratios of these values
        cand = {"B": B, "layout": layout, "levels": levels, "rows": ry, "zchunk": zc,
                "shape": shape,
                "ms_per_step": float(f"{ms:.6g}"), "per_source_gvox": float(f"{per_src:.6g}"),
                "alg_GBps": float(f"{gbs:.6g}")}
        out["candidates"].append(cand)
        log(json.dumps(cand))
        return cand

    with Solver(device) as s:
        s.set_domain_labels(scene.labels)
        s.set_materials(scene.material_table())
        for B in batches:
            load_batch(s, scene, B, rec_per_src)
            s.set_time_levels(1)
            best = None
            for ry in rows:
                for zc in zchunks:
                    s.set_update_tiling(ry, zc)
                    cand = record(B, 1, ry, zc, time_config(s, steps, warmup, N.F32))
                    if best is None or cand["per_source_gvox"] > best["per_source_gvox"]:
                        best = cand
            out["best_single_level"][str(B)] = dict(best)
            if B > 1:
                s.set_batch_layout(N.BATCH_GROUPED)
                for name, shape in SHAPES.items():
                    s.set_sweep_shape(shape)
                    for zc in TWO_LEVEL_ZCHUNKS:
                        s.set_time_levels(2, zc)
                        cand = record(B, 2, None, zc, time_config(s, steps, warmup, N.F32), name,
                                      "grouped")
                        if cand["per_source_gvox"] > best["per_source_gvox"]:
                            best = cand
                s.set_sweep_shape(N.SWEEP_AUTO)
                for zc in THREE_LEVEL_ZCHUNKS:
                    s.set_time_levels(3, zc)
                    cand = record(B, 3, None, zc, time_config(s, steps, warmup, N.F32), "square",
                                  "grouped")
                    if cand["per_source_gvox"] > best["per_source_gvox"]:
                        best = cand
                s.set_time_levels(1)
                s.set_batch_layout(N.BATCH_INTERLEAVED)
            if B == 1:
                for name, shape in SHAPES.items():
                    s.set_sweep_shape(shape)
                    for zc in TWO_LEVEL_ZCHUNKS:
                        s.set_time_levels(2, zc)
                        cand = record(B, 2, None, zc, time_config(s, steps, warmup, N.F32), name)
                        if cand["per_source_gvox"] > best["per_source_gvox"]:
                            best = cand
                s.set_sweep_shape(N.SWEEP_AUTO)
                for zc in THREE_LEVEL_ZCHUNKS:
                    s.set_time_levels(3, zc)
                    cand = record(B, 3, None, zc, time_config(s, steps, warmup, N.F32), "square")
                    if cand["per_source_gvox"] > best["per_source_gvox"]:
                        best = cand
                s.set_time_levels(1)
            out["best"][str(B)] = best
    for table in (out["best"], out["best_single_level"]):
        base = table.get("1")
        if base:
            for v in table.values():
                v["per_source_speedup_vs_B1"] = float(f"{v['per_source_gvox'] / base['per_source_gvox']:.4g}")
    top = max(out["best"].values(), key=lambda c: c["per_source_gvox"])
    out["choice"] = {k: top[k] for k in ("B", "layout", "levels", "rows", "zchunk", "shape")}
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--batches", default="1,8,16,32,64")
    ap.add_argument("--rows", default="1,2,4,8")
    ap.add_argument("--zchunks", default="4,8,16")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=int, default=1, help="divide every room dimension")
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    batches = [int(v) for v in args.batches.split(",")]
    scene = rooms.scene_c4(n_sources=max(batches), n_steps=args.steps, scale=args.scale)
    rep = sweep(scene, batches, [int(v) for v in args.rows.split(",")],
                [int(v) for v in args.zchunks.split(",")], args.steps, args.warmup,
                log=lambda m: print(m, file=sys.stderr, flush=True))
    text = json.dumps(rep)
    if args.out:
        Path(args.out).parent.mkdir(parents=True, exist_ok=True)
        Path(args.out).write_text(text)
    print(json.dumps({"scene": rep["scene"], "best": rep["best"],
                      "best_single_level": rep["best_single_level"], "choice": rep["choice"]}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
