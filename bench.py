#!/usr/bin/env python3
"""Benchmark of the SRL_FORWARD time-stepping core on one MI355X.

Workload (one "step" = one leapfrog time step of the whole room):
  box room 2048 x 1536 x 768 voxels (2,415,919,104 nodes) with 32 seeded solid boxes,
  8 materials (admittance uniform in [0.01, 0.6]), one Gaussian-derivative source,
  64 receivers -- fields and position/material bytes resident in HBM before timing.

Prints one JSON line:
  value     = Gvoxel-updates/s of the FP32 run (all nodes, shell included), host wall clock
              over exactly --steps steps after --warmup untimed steps;
  fp16      = the same with binary16 field storage (power-of-two scaled), and beside it the
              binary16 difference form (P, P - P_past) on rooms of even X;
  roofline  = algorithmic bytes per update-kernel launch (14 B/node FP32) / average update
              kernel duration measured with HIP events on the library's stream, vs 8 TB/s;
              `traffic` = measured HBM bytes per launch from the committed rocprofv3 counter
              summary when one exists, and `achieved_measured` / `frac_measured` = those bytes
              over the same launch time (the rate HBM actually moved);
  shard_share = the same FP32 sweep on one GPU's share of the room under a z-slab
              decomposition over 8 GPUs (2048 x 1536 x 97: 95 updated slices + 2 halo);
  serving   = p50/p99 latency of the continuous-batching pool on a Poisson stream of
              impulse-response jobs of the C5 kind (rooms 64..512 voxels per side);
  cpu_baseline = the CPU oracle (a C restatement of the reference update loop) timed on this
              host on a z-slab sample of the same room, on every host thread this process may
              use (OMP_NUM_THREADS caps it where set), with single-thread, literal-mode and
              130x98x66 full-run figures beside it.

`--bw-probe` prints the part's streaming ceilings instead (read, write, copy, triad, in-place
update over three C3-sized arrays), measured by the library's own probe kernels.
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
# HBM bytes per launch of the dominant FP32 kernel (three time levels per launch), from counters
TRAFFIC_SUMMARY = ROOT / "profiles" / "srl_update3_c3_f32_hbm_traffic.md"
# steps of the kernel-timing runs: whole launches of the two- and the three-level sweeps
KERNEL_STEPS = 96


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


def host_threads() -> dict:
    """Host threads this process may use: `nproc` (the affinity mask), the machine's online
    processors (what std::thread::hardware_concurrency reports, sysconf _SC_NPROCESSORS_ONLN)
    and the OpenMP cap the environment sets, if any."""
    nproc = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    cap = int(omp) if omp.isdigit() and int(omp) > 0 else None
    return {"nproc": nproc, "hardware_concurrency": os.cpu_count(), "omp_num_threads": cap,
            "used": min(nproc, cap) if cap else nproc}


def _oracle_rate(pos, mat, table, src, sig, rec, steps, threads, literal=0) -> tuple[float, float]:
    import oracle

    t0 = time.perf_counter()
    oracle.run(pos, mat, table, src, np.array([1], np.int32), sig, rec, steps, threads=threads,
               literal=literal)
    dt = time.perf_counter() - t0
    return pos.size * steps / dt / 1e9, dt


def cpu_baseline(scene, steps: int = 300, nz: int = 66) -> dict:
    """Oracle timed on a z-slab of the same room (the first nz-1 slices plus a closing shell):
    every usable host thread (the reported value), one thread (a 10-slice slab), and the
    reference's literal material slot (kernels3d.cu:513) on every thread; plus the 130x98x66
    box run in full (2000 steps) on the CPU and on the GPU."""
    import oracle

    lab = np.ascontiguousarray(scene.labels[:nz].copy())
    lab[nz - 1] = 2  # ceiling material closes the slab
    pos, mat, _ = oracle.classify(lab)
    pts = rooms.seeded_points(lab, rooms.SplitMix64(99), 2)
    ht = host_threads()
    threads = ht["used"]
    table = scene.material_table()
    sig = scene.signals[:, :steps].astype(np.float32)
    rate, dt = _oracle_rate(pos, mat, table, pts[:1], sig, pts[1:], steps, threads)
    # literal slot: every node reads table slot mat*20*octave = 0 at octave 0 (same cost)
    lit_steps = max(1, steps // 3)
    rate_lit, dt_lit = _oracle_rate(pos, mat, table, pts[:1], sig[:, :lit_steps], pts[1:],
                                    lit_steps, threads, literal=1)
    one = np.ascontiguousarray(lab[:11].copy())
    one[10] = 2
    pos1, mat1, _ = oracle.classify(one)
    p1 = rooms.seeded_points(one, rooms.SplitMix64(98), 2)
    one_steps = 40
    rate1, dt1 = _oracle_rate(pos1, mat1, table, p1[:1], sig[:, :one_steps], p1[1:], one_steps, 1)
    Z, Y, X = lab.shape
    c1 = rooms.scene_c1(n_steps=2000)
    pc, mc, _ = oracle.classify(c1.labels)
    t0 = time.perf_counter()
    oracle.run(pc, mc, c1.material_table(), c1.src_xyz, c1.src_kind,
               c1.signals.astype(np.float32), c1.rec_xyz, 2000, threads=threads)
    c1_cpu = time.perf_counter() - t0
    with Solver(0) as s:
        s.load_scene(c1)
        s.run(20, N.F32)
        t0 = time.perf_counter()
        s.run(2000, N.F32)
        c1_gpu = time.perf_counter() - t0
    return {"value": round(rate, 4), "unit": "Gvoxel-updates/s",
            "cores": threads, "kind": "port",
            "sample": f"{X}x{Y}x{Z} z-slab of the same room, {steps} FP32 steps, "
                      f"OpenMP over z on {threads} threads, {dt:.1f} s",
            "host": ht,
            "single_thread": {"value": round(rate1, 4), "sample":
                              f"{X}x{Y}x11 slab, {one_steps} steps, 1 thread, {dt1:.1f} s"},
            "literal_mode": {"value": round(rate_lit, 4), "sample":
                             f"the {X}x{Y}x{Z} slab, {lit_steps} steps, literal material slot "
                             f"(kernels3d.cu:513), {threads} threads, {dt_lit:.1f} s"},
            "c1_full_run": {"workload": "130x98x66 box, 2000 FP32 steps, 1 soft source, "
                                        "1 receiver",
                            "cpu_s": round(c1_cpu, 3), "gpu_s": round(c1_gpu, 4),
                            "cpu_threads": threads,
                            "cpu_gvox_per_s": round(c1.n_nodes * 2000 / c1_cpu / 1e9, 4),
                            "gpu_gvox_per_s": round(c1.n_nodes * 2000 / c1_gpu / 1e9, 3)}}


def shard_share(scene, steps: int, warmup: int, n_gpus: int = 8) -> dict:
    """One GPU's share of the room under a z-slab decomposition over n_gpus GPUs: Z/n_gpus
    updated slices plus the two halo slices, here closed by the ceiling material (the update
    work is the same), FP32, the default sweep (three levels per launch)."""
    Z = scene.labels.shape[0]
    nz = (Z - 2) // n_gpus + 2
    lab = np.ascontiguousarray(scene.labels[:nz].copy())
    lab[nz - 1] = 2
    pts = rooms.seeded_points(lab, rooms.SplitMix64(97), 9)
    sc = rooms.Scene(f"{scene.name}_share{n_gpus}", lab, scene.betas, pts[:1],
                     np.array([1], np.int32), scene.signals[:1], pts[1:], scene.n_steps)
    with Solver(0) as s:
        s.load_scene(sc)
        wall, _ = timed_run(s, steps, warmup, N.F32)
        kms = kernel_avg_ms(s, KERNEL_STEPS, N.F32)
    n = lab.size
    return {"workload": f"z-slab of the room for one of {n_gpus} GPUs: "
                        f"{lab.shape[2]}x{lab.shape[1]}x{nz} ({nz - 2} updated slices + 2 halo)",
            "nodes": int(n), "value": round(n * steps / wall / 1e9, 2),
            "ms_per_step": round(wall * 1e3 / steps, 4), "kernel_ms": round(kms, 4),
            "roofline_frac": round(n * BYTES_PER_NODE[N.F32] / (kms * 1e-3) / 1e9 / PEAK_GBS, 4)}


def bw_probe(bytes_per_array: int, iters: int) -> dict:
    """Streaming ceilings of this part from the library's probe kernels."""
    import ctypes as C

    out = {}
    for name, kind in (("read", N.PROBE_READ), ("write", N.PROBE_WRITE), ("copy", N.PROBE_COPY),
                       ("triad", N.PROBE_TRIAD), ("update_in_place", N.PROBE_UPDATE)):
        ms, moved = C.c_double(), C.c_double()
        N.check(N.lib().pfdtd_hbm_probe(0, kind, bytes_per_array, iters, C.byref(ms),
                                        C.byref(moved)), "pfdtd_hbm_probe")
        gbs = moved.value / (ms.value * 1e-3) / 1e9
        out[name] = {"ms_per_launch": round(ms.value, 4), "bytes_per_launch": moved.value,
                     "GB_per_s": round(gbs, 1), "frac_of_8TBs": round(gbs / PEAK_GBS, 4)}
        log(f"{name}: {gbs:.0f} GB/s ({ms.value:.3f} ms per launch)")
    return out


SERVE_RATE = 2.5  # jobs/s: below the pool's capacity on this stream (about 3.9 jobs/s)


def serving_leg(jobs: int, rate: float, scale: int = 1) -> dict:
    """p50/p99 impulse-response latency of the continuous-batching pool on a Poisson stream."""
    from parallelfdtd_amd import serve

    rep = serve.serve(jobs, rate, scale=scale, step_scale=1, log=log)
    size = "" if scale == 1 else f" at 1/{scale} size"
    return {"jobs": jobs, "rate_jobs_per_s": rate,
            "stream": f"C5 impulse-response stream{size}: the first {jobs} jobs, rooms "
                      f"{64 // scale}..{512 // scale} voxels per side, 0-10 boxes, 1-6 materials, "
                      f"1-16 sources, 2000-8000 steps, FP32/FP16 1:1, Poisson arrivals",
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
    ap.add_argument("--fp64-only", action="store_true",
                    help="only the FP64 run (three-level sweep; profiling); prints no bench line")
    ap.add_argument("--f16diff-only", action="store_true",
                    help="only the FP16 difference-form run (profiling); prints no bench line")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 4 GiB device copy (counter calibration for rocprofv3)")
    ap.add_argument("--bw-probe", action="store_true",
                    help="only measure the HBM streaming ceilings (read/write/copy/triad/update)")
    ap.add_argument("--probe-gib", type=float, default=9.0,
                    help="bytes per probe array in GiB (the C3 FP32 field is 9.0 GiB)")
    args = ap.parse_args()

    if args.bw_probe:
        rep = {"probe": "pfdtd_hbm_probe (float4 per lane, 4 in flight)",
               "bytes_per_array": int(args.probe_gib * 2**30), "iters": 10,
               "results": bw_probe(int(args.probe_gib * 2**30), 10)}
        print(json.dumps(rep), flush=True)
        return 0

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

        if args.fp16_only or args.f16diff_only or args.fp64_only:
            prec, name = ((N.F16, "FP16 storage") if args.fp16_only else
                          (N.F64, "FP64") if args.fp64_only else
                          (N.F16_DIFF, "FP16 difference form"))
            if prec != N.F64:
                s.prepare_fields(prec, 12)
            wall16, _ = timed_run(s, args.steps, args.warmup, prec)
            k16 = kernel_avg_ms(s, KERNEL_STEPS, prec)
            log(f"{name}: {nodes * args.steps / wall16 / 1e9:.1f} Gvox/s, "
                f"update {k16:.3f} ms/step")
            return 0
        wall, dev_ms = timed_run(s, args.steps, args.warmup, N.F32)
        value = nodes * args.steps / wall / 1e9
        log(f"FP32: {value:.1f} Gvox/s wall, device {nodes * args.steps / dev_ms / 1e6:.1f}")
        kms = kernel_avg_ms(s, KERNEL_STEPS, N.F32)
        achieved = nodes * BYTES_PER_NODE[N.F32] / (kms * 1e-3) / 1e9
        log(f"FP32 update kernel: {kms:.3f} ms/step, {achieved:.0f} GB/s algorithmic")

        fp16 = None
        if not args.fp32_only:
            s.prepare_fields(N.F16, 12)
            wall16, _ = timed_run(s, args.steps, args.warmup, N.F16)
            k16 = kernel_avg_ms(s, KERNEL_STEPS, N.F16)
            fp16 = {"value": round(nodes * args.steps / wall16 / 1e9, 2),
                    "ms_per_step": round(wall16 * 1e3 / args.steps, 4),
                    "kernel_ms": round(k16, 4),
                    "roofline_frac": round(nodes * 8 / (k16 * 1e-3) / 1e9 / PEAK_GBS, 4),
                    "speedup_vs_fp32": round(wall / wall16, 3)}
            log(f"FP16 storage: {fp16}")
            if X % 2 == 0:
                # binary16 difference form (P, P - P_past): the two-column sweep, 3 waves/SIMD
                s.prepare_fields(N.F16_DIFF, 12)
                walld, _ = timed_run(s, args.steps, args.warmup, N.F16_DIFF)
                kd = kernel_avg_ms(s, KERNEL_STEPS, N.F16_DIFF)
                fp16["difference_form"] = {
                    "value": round(nodes * args.steps / walld / 1e9, 2),
                    "ms_per_step": round(walld * 1e3 / args.steps, 4),
                    "kernel_ms": round(kd, 4),
                    "roofline_frac": round(nodes * 8 / (kd * 1e-3) / 1e9 / PEAK_GBS, 4),
                    "speedup_vs_fp32": round(wall / walld, 3)}
                log(f"FP16 difference form: {fp16['difference_form']}")

    traffic = committed_traffic() if args.scale == 1 else None
    launch_s = 3 * kms * 1e-3
    measured = traffic / launch_s / 1e9 if traffic else None
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
                     "achieved_kind": "algorithmic bytes (14 B per node-update, 3 updates per "
                                      "launch) / launch time; above 1 of peak because a launch "
                                      "moves each field once for three steps",
                     "achieved_measured": round(measured, 1) if measured else None,
                     "frac_measured": round(measured / PEAK_GBS, 4) if measured else None,
                     "kernel": "srl_update3_kernel<float> (three time levels per launch)",
                     "launch_ms": round(3 * kms, 4), "ms_per_step": round(kms, 4),
                     "bytes_per_node_update": BYTES_PER_NODE[N.F32]},
        "fp16_storage": fp16,
    }
    if not args.fp32_only and args.scale == 1:
        out["shard_share"] = shard_share(sc, args.steps, args.warmup)
        log(f"shard share: {out['shard_share']}")
    if args.serve_jobs > 0:
        out["serving"] = serving_leg(args.serve_jobs, args.serve_rate, args.scale)
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
