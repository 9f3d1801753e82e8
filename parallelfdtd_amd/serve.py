"""Serving loop: a stream of impulse-response jobs (rooms of different sizes, materials, sources,
run lengths and storage precisions) served by one device-resident job pool with continuous
batching.

The pool (``pfdtd_pool_*`` in libpfdtd.so) keeps every admitted job's fields resident in one HBM
arena and advances all resident jobs together, one multi-job kernel launch per step and storage
type, in hipGraph-captured bundles of steps. This module supplies the request side: it admits
jobs first-come-first-served as they arrive, whenever the pool has arena space and a slot free
(a job that does not fit waits at the head of the queue), keeps bundles flowing while jobs are
resident, collects results as jobs finish, and reports latency (completion - arrival) and the
aggregate voxel-update rate.

The synthetic stream (make_job): job j uses SplitMix64(5).fork(j); each room dimension uniform in
64..512 voxels (divided by `scale`), 0-10 solid boxes, 1-6 materials with admittance uniform in
[0.01, 0.6], 1-16 soft sources at K=6 points driven by a unit Gaussian pulse (sigma 10 steps,
centre 50), 4 receivers, 2000-8000 steps (divided by `step_scale`), FP32 or FP16 storage with
probability 1/2 each (FP16 scaled by 2^12). Arrivals are a Poisson process of the stated rate.

Run:  python -m parallelfdtd_amd.serve --jobs 1000 --rate 2.0 --out serve_c5.json
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import sys
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import native as N
from . import rooms
from .rooms import SplitMix64

LAMBDA = math.sqrt(1.0 / 3.0)


@dataclass
class JobSpec:
    """Inputs of one impulse-response job (the pfdtd_job fields)."""
    labels: np.ndarray          # (Z, Y, X) uint8
    betas: np.ndarray           # (n_mat,) float32 admittances
    src_xyz: np.ndarray         # (n_src, 3) int32
    src_kind: np.ndarray        # (n_src,) int32
    signals: np.ndarray         # (n_src, n_samples) float32
    rec_xyz: np.ndarray         # (n_rec, 3) int32
    n_steps: int
    precision: int = N.F32
    scale_exp: int = 0
    flags: int = 0
    octave: int = 0
    lam: float = LAMBDA

    @property
    def voxels(self) -> int:
        return int(self.labels.size)

    def struct(self) -> tuple[N.PfdtdJob, list]:
        """pfdtd_job view of the spec (plus the arrays it points into, to keep alive)."""
        lab = np.ascontiguousarray(self.labels, np.uint8)
        mats = np.ascontiguousarray(np.repeat(self.betas.astype(np.float32)[:, None], 20, axis=1))
        sx = np.ascontiguousarray(self.src_xyz, np.int32).reshape(-1, 3)
        sk = np.ascontiguousarray(self.src_kind, np.int32).reshape(-1)
        sg = np.ascontiguousarray(self.signals, np.float32).reshape(len(sk), -1)
        rx = np.ascontiguousarray(self.rec_xyz, np.int32).reshape(-1, 3)
        Z, Y, X = lab.shape
        j = N.PfdtdJob(labels=lab.ctypes.data, X=X, Y=Y, Z=Z, materials=mats.ctypes.data,
                       n_mat=mats.shape[0], n_coef=20, octave=self.octave, lambda_=self.lam,
                       flags=self.flags, n_src=len(sk), src_xyz=sx.ctypes.data,
                       src_kind=sk.ctypes.data, n_samples=sg.shape[1], signals=sg.ctypes.data,
                       n_rec=len(rx), rec_xyz=rx.ctypes.data, n_steps=self.n_steps,
                       precision=self.precision, scale_exp=self.scale_exp)
        return j, [lab, mats, sx, sk, sg, rx]


def make_job(job_id: int, scale: int = 1, step_scale: int = 1) -> JobSpec:
    """Job `job_id` of the synthetic stream (see the module docstring)."""
    rng = SplitMix64(5).fork(job_id)
    X, Y, Z = (max(8, rng.randint(64, 512) // scale) for _ in range(3))
    n_mat = rng.randint(1, 6)
    betas = np.array([rng.uniform(0.01, 0.6) for _ in range(n_mat)], np.float32)
    lab = rooms.box_room(X, Y, Z, floor=0, ceiling=min(1, n_mat - 1), walls=min(2, n_mat - 1))
    n_box = rng.randint(0, 10)
    smax = max(2, min(X, Y, Z) // 4)
    rooms.add_boxes(lab, rng.fork(3), n_box, 2, smax, list(range(n_mat)))
    n_src = rng.randint(1, 16)
    src = rooms.seeded_points(lab, rng.fork(4), n_src)
    rec = rooms.seeded_points(lab, rng.fork(5), 4)
    n_steps = max(16, rng.randint(2000, 8000) // step_scale)
    fp16 = rng.uniform() < 0.5
    sig = np.repeat(rooms.gaussian_pulse(min(n_steps, 200), 50.0, 10.0)[None, :], n_src, axis=0)
    return JobSpec(lab, betas, src, np.ones(n_src, np.int32), sig.astype(np.float32), rec,
                   n_steps, N.F16 if fp16 else N.F32, 12 if fp16 else 0)


def run_alone(spec: JobSpec, device: int = 0) -> np.ndarray:
    """The job on its own context (pfdtd_run): the reference result for the pool's answer."""
    from .solver import Solver
    with Solver(device) as s:
        s.set_domain_labels(spec.labels)
        s.set_materials(np.repeat(spec.betas[:, None], 20, axis=1), octave=spec.octave,
                        lam=spec.lam, literal_slot=bool(spec.flags & N.MATSLOT_LITERAL),
                        soft_overwrite=bool(spec.flags & N.SOFT_OVERWRITE))
        s.set_sources(spec.src_xyz, spec.src_kind, signal_f32=spec.signals)
        s.set_receivers(spec.rec_xyz)
        if spec.precision in (N.F16, N.BF16):
            s.prepare_fields(spec.precision, spec.scale_exp)
        s.run(spec.n_steps, spec.precision)
        return s.responses(np.float32)


class Pool:
    """Thin wrapper of the pfdtd_pool C ABI."""

    def __init__(self, device: int = 0, arena_bytes: int = 0, bundle: int = 32,
                 use_graph: bool = True, rows: int = 2, zchunk: int = 8, max_jobs: int = 64):
        self._lib = N.lib()
        self._h = C.c_void_p()
        N.check(self._lib.pfdtd_pool_create(device, int(arena_bytes), C.byref(self._h)),
                "pfdtd_pool_create")
        self._c("pfdtd_pool_set_options", bundle, 1 if use_graph else 0, rows, zchunk, max_jobs)

    def _c(self, name: str, *args) -> None:
        N.check(getattr(self._lib, name)(self._h, *args), name)

    def close(self) -> None:
        if self._h:
            self._lib.pfdtd_pool_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def job_bytes(spec: JobSpec) -> int:
        j, _keep = spec.struct()
        b = C.c_int64()
        N.check(N.lib().pfdtd_pool_job_bytes(C.byref(j), C.byref(b)), "pfdtd_pool_job_bytes")
        return b.value

    def mem(self) -> tuple[int, int, int]:
        a, f, l = C.c_int64(), C.c_int64(), C.c_int64()
        self._c("pfdtd_pool_mem", C.byref(a), C.byref(f), C.byref(l))
        return a.value, f.value, l.value

    def submit(self, spec: JobSpec) -> int | None:
        """Job id, or None when the job does not fit now (arena space or slots)."""
        j, _keep = spec.struct()
        jid = C.c_int64()
        rc = self._lib.pfdtd_pool_submit(self._h, C.byref(j), C.byref(jid))
        if rc == N.E_NOMEM:
            return None
        N.check(rc, "pfdtd_pool_submit")
        return jid.value

    def advance(self) -> int:
        n = C.c_int()
        self._c("pfdtd_pool_advance", C.byref(n))
        return n.value

    def poll(self) -> list[int]:
        ids = (C.c_int64 * 64)()
        out: list[int] = []
        while True:
            n = C.c_int()
            self._c("pfdtd_pool_poll", ids, 64, C.byref(n))
            out.extend(ids[i] for i in range(n.value))
            if n.value < 64:
                return out

    def drain(self) -> None:
        self._c("pfdtd_pool_drain")

    def info(self, jid: int) -> tuple[int, int, int]:
        ph, nr, ns = C.c_int(), C.c_int(), C.c_int()
        self._c("pfdtd_pool_job_info", jid, C.byref(ph), C.byref(nr), C.byref(ns))
        return ph.value, nr.value, ns.value

    def responses(self, jid: int, dtype=np.float32) -> np.ndarray:
        _, nr, ns = self.info(jid)
        out = np.zeros((nr, ns), dtype)
        if dtype == np.float32:
            self._c("pfdtd_pool_get_responses", jid, N.ptr(out), None, out.size)
        else:
            self._c("pfdtd_pool_get_responses", jid, None, N.ptr(out), out.size)
        return out

    def release(self, jid: int) -> None:
        self._c("pfdtd_pool_release", jid)

    def stats(self) -> dict:
        b, g, v, r = C.c_int64(), C.c_int64(), C.c_double(), C.c_int()
        self._c("pfdtd_pool_stats", C.byref(b), C.byref(g), C.byref(v), C.byref(r))
        return {"bundles": b.value, "graph_builds": g.value, "voxel_steps": v.value,
                "resident": r.value}


def poisson_arrivals(n: int, rate: float, seed: int = 5) -> np.ndarray:
    """Arrival times (s) of a Poisson process of `rate` jobs/s."""
    rng = SplitMix64(seed).fork(0xA7)
    t, out = 0.0, []
    for _ in range(n):
        t += -math.log(1.0 - rng.uniform()) / rate
        out.append(t)
    return np.asarray(out)


def serve(n_jobs: int, rate: float, scale: int = 1, step_scale: int = 1, arena_bytes: int = 0,
          bundle: int = 32, use_graph: bool = True, max_jobs: int = 64, keep_results: bool = False,
          log=print, progress_every: float = 10.0) -> dict:
    """Serve the synthetic stream; returns latency / throughput statistics (and, with
    keep_results, every job's responses)."""
    arrivals = poisson_arrivals(n_jobs, rate)
    results: dict[int, np.ndarray] = {}
    rec: dict[int, dict] = {}
    by_pool_id: dict[int, int] = {}
    nxt = 0           # next job to arrive
    queue: list[int] = []  # arrived, waiting for admission (FIFO)
    pending_spec: dict[int, JobSpec] = {}
    done = 0
    with Pool(0, arena_bytes, bundle, use_graph, max_jobs=max_jobs) as pool:
        arena = pool.mem()[0]
        t0 = time.perf_counter()
        last_log = t0
        active = 0
        while done < n_jobs:
            now = time.perf_counter() - t0
            while nxt < n_jobs and arrivals[nxt] <= now:
                queue.append(nxt)
                nxt += 1
            while queue:
                j = queue[0]
                spec = pending_spec.get(j) or make_job(j, scale, step_scale)
                jid = pool.submit(spec)
                if jid is None:
                    if not by_pool_id:
                        raise RuntimeError(f"job {j} ({Pool.job_bytes(spec)} bytes) does not "
                                           f"fit in an empty pool of {arena} bytes")
                    pending_spec[j] = spec  # keep the built room for the next try
                    break
                pending_spec.pop(j, None)
                queue.pop(0)
                by_pool_id[jid] = j
                rec[j] = {"arrival": float(arrivals[j]), "admit": time.perf_counter() - t0,
                          "voxels": spec.voxels, "steps": spec.n_steps,
                          "precision": "f16" if spec.precision == N.F16 else "f32"}
            active = pool.advance()
            for jid in pool.poll():
                j = by_pool_id.pop(jid)
                rec[j]["done"] = time.perf_counter() - t0
                if keep_results:
                    results[j] = pool.responses(jid)
                pool.release(jid)
                done += 1
            if active == 0 and not queue and nxt < n_jobs and not by_pool_id:
                time.sleep(max(0.0, min(0.05, arrivals[nxt] - (time.perf_counter() - t0))))
            elif active == 0:
                time.sleep(2e-4)  # last bundles of finishing jobs still in flight
            if time.perf_counter() - last_log > progress_every:
                last_log = time.perf_counter()
                log(f"[serve] t={last_log - t0:.1f}s arrived {nxt} done {done} resident "
                    f"{len(by_pool_id)} queued {len(queue)}")
        t_end = time.perf_counter() - t0
        st = pool.stats()
    lat = np.array([rec[j]["done"] - rec[j]["arrival"] for j in range(n_jobs)])
    wait = np.array([rec[j]["admit"] - rec[j]["arrival"] for j in range(n_jobs)])
    vox_steps = float(sum(rec[j]["voxels"] * rec[j]["steps"] for j in range(n_jobs)))
    span = t_end - float(arrivals[0])
    out = {"jobs": n_jobs, "rate_jobs_per_s": rate, "scale": scale, "step_scale": step_scale,
           "bundle_steps": bundle, "graph": bool(use_graph), "arena_bytes": arena,
           "latency_s": {"p50": float(np.percentile(lat, 50)), "p90": float(np.percentile(lat, 90)),
                         "p99": float(np.percentile(lat, 99)), "mean": float(lat.mean()),
                         "max": float(lat.max())},
           "queue_wait_s": {"p50": float(np.percentile(wait, 50)),
                            "p99": float(np.percentile(wait, 99))},
           "voxel_updates": vox_steps, "makespan_s": span,
           "gvox_per_s": vox_steps / span / 1e9, "pool": st,
           "fp16_jobs": int(sum(r["precision"] == "f16" for r in rec.values()))}
    if keep_results:
        out["results"] = results
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--jobs", type=int, default=1000)
    ap.add_argument("--rate", type=float, default=2.0, help="Poisson arrival rate, jobs/s")
    ap.add_argument("--scale", type=int, default=1, help="divide every room dimension")
    ap.add_argument("--step-scale", type=int, default=1, help="divide every run length")
    ap.add_argument("--arena-gb", type=float, default=0.0, help="pool arena (0: 7/8 of free HBM)")
    ap.add_argument("--bundle", type=int, default=32)
    ap.add_argument("--no-graph", action="store_true", help="direct launches instead of hipGraph")
    ap.add_argument("--max-jobs", type=int, default=64)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    rep = serve(args.jobs, args.rate, args.scale, args.step_scale, int(args.arena_gb * 2**30),
                args.bundle, not args.no_graph, args.max_jobs,
                log=lambda m: print(m, file=sys.stderr, flush=True))
    text = json.dumps(rep)
    if args.out:
        Path(args.out).parent.mkdir(parents=True, exist_ok=True)
        Path(args.out).write_text(text)
    print(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
