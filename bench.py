#!/usr/bin/env python3
"""Benchmark of the SRL_FORWARD time-stepping core on one MI355X.

Workload (one "step" = one leapfrog time step of the whole room):
  box room 2048 x 1536 x 768 voxels (2,415,919,104 nodes) with 32 seeded solid boxes,
  8 materials (admittance uniform in [0.01, 0.6]), one Gaussian-derivative source,
  64 receivers -- fields and position/material bytes resident in HBM before timing.

Prints one JSON line:
  value     = Gvoxel-updates/s of the FP32 run (all nodes, shell included), host wall clock
              over exactly --steps steps after --warmup untimed steps;
  fp16      = the same with binary16 field storage (power-of-two scaled);
  roofline  = algorithmic bytes per update-kernel launch (14 B/node FP32) / average update
              kernel duration measured with HIP events on the library's stream, vs 8 TB/s;
              `traffic` = measured HBM bytes per launch from the committed rocprofv3 counter
              summary when one exists;
  cpu_baseline = the CPU oracle (a C restatement of the reference update loop) timed on this
              host on a z-slab sample of the same room.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from parallelfdtd_amd import native as N  # noqa: E402
from parallelfdtd_amd import rooms  # noqa: E402
from parallelfdtd_amd.solver import Solver  # noqa: E402

METRIC = "Gvoxel-updates/s + % HBM roofline (FP32 vs FP16-store); p50/p99 IR latency"
PEAK_GBS = 8000.0
BYTES_PER_NODE = {N.F32: 14, N.F16: 8, N.BF16: 8, N.F64: 26}
# HBM bytes per launch of the dominant FP32 kernel (two time levels per launch), from counters
TRAFFIC_SUMMARY = ROOT / "profiles" / "srl_update2_c3_f32_hbm_traffic.md"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def timed_run(s: Solver, steps: int, warmup: int, prec: int) -> tuple[float, float]:
    """(wall seconds, device ms) of exactly `steps` steps after `warmup` untimed steps."""
    if warmup:
        s.run(warmup, prec)
    t0 = time.perf_counter()
    done, dev_ms = s.run(steps, prec)
    wall = time.perf_counter() - t0
    assert done == steps
    return wall, dev_ms


def kernel_avg_ms(s: Solver, steps: int, prec: int) -> float:
    s.set_kernel_timing(True)
    s.run(steps, prec)
    ms, n = s.kernel_timing()
    s.set_kernel_timing(False)
    return ms / n


def committed_traffic() -> float | None:
    """HBM bytes per update launch from the committed counter summary, if present."""
    if not TRAFFIC_SUMMARY.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", TRAFFIC_SUMMARY.read_text())
    return float(m.group(1)) if m else None


def cpu_baseline(scene, steps: int = 300, nz: int = 66) -> dict:
    """Oracle timed on a z-slab of the same room: the first nz-1 slices plus a closing shell."""
    import oracle

    lab = np.ascontiguousarray(scene.labels[:nz].copy())
    lab[nz - 1] = 2  # ceiling material closes the slab
    pos, mat, _ = oracle.classify(lab)
    pts = rooms.seeded_points(lab, rooms.SplitMix64(99), 2)
    threads = min(16, len(os.sched_getaffinity(0)))
    sig = scene.signals[:, :steps].astype(np.float32)
    t0 = time.perf_counter()
    oracle.run(pos, mat, scene.material_table(), pts[:1], np.array([1], np.int32), sig, pts[1:],
               steps, threads=threads)
    dt = time.perf_counter() - t0
    Z, Y, X = lab.shape
    return {"value": round(lab.size * steps / dt / 1e9, 4), "unit": "Gvoxel-updates/s",
            "cores": threads, "kind": "port",
            "sample": f"{X}x{Y}x{Z} z-slab of the same room, {steps} FP32 steps, "
                      f"OpenMP over z on {threads} threads, {dt:.1f} s"}


SERVE_SCALE = 2  # serving leg: rooms 32..256 voxels per side (the C5 stream at half size)
SERVE_RATE = 20.0


def serving_leg(jobs: int, rate: float) -> dict:
    """p50/p99 impulse-response latency of the continuous-batching pool on a Poisson stream."""
    from parallelfdtd_amd import serve

    rep = serve.serve(jobs, rate, scale=SERVE_SCALE, step_scale=1, log=log)
    return {"jobs": jobs, "rate_jobs_per_s": rate,
            "stream": f"C5 stream at 1/{SERVE_SCALE} size: rooms {64 // SERVE_SCALE}.."
                      f"{512 // SERVE_SCALE} per side, 0-10 boxes, 1-6 materials, 1-16 sources, "
                      f"2000-8000 steps, FP32/FP16 1:1",
            "p50_latency_s": round(rep["latency_s"]["p50"], 4),
            "p99_latency_s": round(rep["latency_s"]["p99"], 4),
            "gvox_per_s": round(rep["gvox_per_s"], 2), "makespan_s": round(rep["makespan_s"], 2),
            "graph_builds": rep["pool"]["graph_builds"], "bundles": rep["pool"]["bundles"]}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--scale", type=int, default=1, help="divide every room dimension (quick runs)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--serve-jobs", type=int, default=200,
                    help="impulse-response jobs in the serving leg (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=SERVE_RATE,
                    help="Poisson arrival rate of the serving leg, jobs/s")
    ap.add_argument("--fp32-only", action="store_true", help="only the FP32 run (profiling)")
    ap.add_argument("--fp16-only", action="store_true",
                    help="only the FP16-storage run (profiling); prints no bench line")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 4 GiB device copy (counter calibration for rocprofv3)")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch

        src = torch.ones(1 << 30, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4}))
        return 0

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the SRL core holds the whole room on one MI355X "
              "(multi-GPU z-slab sharding is a design note, not part of this build)", file=sys.stderr)
        return 2

    t0 = time.perf_counter()
    sc = rooms.scene_c3(n_steps=max(args.steps, args.warmup, 1), scale=args.scale)
    X, Y, Z = sc.shape_xyz
    nodes = sc.n_nodes
    log(f"scene {sc.name}: {nodes} nodes, built in {time.perf_counter() - t0:.1f} s")

    out: dict = {}
    with Solver(0) as s:
        t0 = time.perf_counter()
        s.load_scene(sc)
        log(f"domain classified on GPU in {time.perf_counter() - t0:.1f} s")

        if args.fp16_only:
            s.prepare_fields(N.F16, 12)
            wall16, _ = timed_run(s, args.steps, args.warmup, N.F16)
            k16 = kernel_avg_ms(s, min(args.steps, 100), N.F16)
            log(f"FP16 storage: {nodes * args.steps / wall16 / 1e9:.1f} Gvox/s, "
                f"update {k16:.3f} ms/step")
            return 0
        wall, dev_ms = timed_run(s, args.steps, args.warmup, N.F32)
        value = nodes * args.steps / wall / 1e9
        log(f"FP32: {value:.1f} Gvox/s wall, device {nodes * args.steps / dev_ms / 1e6:.1f}")
        kms = kernel_avg_ms(s, min(args.steps, 100), N.F32)
        achieved = nodes * BYTES_PER_NODE[N.F32] / (kms * 1e-3) / 1e9
        log(f"FP32 update kernel: {kms:.3f} ms/launch, {achieved:.0f} GB/s algorithmic")

        fp16 = None
        if not args.fp32_only:
            s.prepare_fields(N.F16, 12)
            wall16, _ = timed_run(s, args.steps, args.warmup, N.F16)
            k16 = kernel_avg_ms(s, min(args.steps, 100), N.F16)
            fp16 = {"value": round(nodes * args.steps / wall16 / 1e9, 2),
                    "ms_per_step": round(wall16 * 1e3 / args.steps, 4),
                    "kernel_ms": round(k16, 4),
                    "roofline_frac": round(nodes * 8 / (k16 * 1e-3) / 1e9 / PEAK_GBS, 4),
                    "speedup_vs_fp32": round(wall / wall16, 3)}
            log(f"FP16 storage: {fp16}")

    traffic = committed_traffic() if args.scale == 1 else None
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Gvoxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"SRL FP32 box room {X}x{Y}x{Z} ({nodes} nodes), 32 boxes, "
                               f"8 materials, 1 source, 64 receivers",
                   "nodes": nodes, "precision": "fp32", "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": "srl_update2_kernel<float> (two time levels per launch)",
                     "launch_ms": round(2 * kms, 4), "ms_per_step": round(kms, 4),
                     "bytes_per_node_update": BYTES_PER_NODE[N.F32]},
        "fp16_storage": fp16,
    }
    if args.serve_jobs > 0:
        out["serving"] = serving_leg(args.serve_jobs, args.serve_rate)
        log(f"serving: {out['serving']}")
    if not args.no_cpu:
        t0 = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline(sc)
        log(f"CPU baseline in {time.perf_counter() - t0:.1f} s: {out['cpu_baseline']}")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
