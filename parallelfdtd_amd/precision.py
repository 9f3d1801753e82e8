"""Precision study: the same room advanced in FP64 and in lower-precision field storage, with
every receiver response reported as relative L2 and peak error against the FP64 run.

Storage formats: FP32 fields and arithmetic; FP16 and BF16 fields with FP32 arithmetic and a
power-of-two scale 2^e chosen so the expected field peak maps to about 2^12 (the source peak
sets it: a unit pulse gives e = 12); and `f16diff`, binary16 storage of P^n and
D^n = P^n - P^{n-1} (the difference form, PFDTD_F16_DIFF; rooms of even X). Errors per receiver r:
  rel_l2[r] = |x_r - ref_r|_2 / |ref_r|_2,   peak[r] = max |x_r - ref_r| / max |ref_r|.
The summary statistics cover the receivers the wave has reached within the run, those whose
FP64 peak is at least REACHED (1e-4) times the largest receiver peak. A receiver outside the
arrival front holds only the stencil's numerical precursor (values down to 1e-300 in FP64,
which FP32/FP16 flush to zero, giving rel_l2 = 1); such receivers are counted and listed, and
the per-receiver arrays keep every receiver. `post_filtered` repeats the statistics after the
reference's response post-filter (FDTDpostFilter.m, low-pass at 0.2 of Nyquist + DC block),
which removes the band the leapfrog does not resolve.

Run:  python -m parallelfdtd_amd.precision --scene c3 --steps 3000 --out reports/precision_c3.json
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

from . import native as N
from . import rooms
from .postfilter import post_filter
from .solver import Solver

# The reference's response conditioning (matlab/testBench.m:204): FDTDpostFilter at 0.2 of the
# Nyquist rate; the 5 Hz DC blocker needs a sampling rate, taken as 40 kHz here.
POST_FS = 40000.0
POST_FRAC = 0.2

FORMATS = {"f32": N.F32, "f16": N.F16, "f16diff": N.F16_DIFF, "bf16": N.BF16}


def scale_exponent(peak: float, target_log2: int = 12) -> int:
    """Power-of-two exponent e so that peak * 2^e is about 2^target_log2."""
    if peak <= 0:
        return 0
    return int(target_log2 - math.floor(math.log2(peak)))


REACHED = 1e-4


def errors(x: np.ndarray, ref: np.ndarray) -> dict:
    num = np.linalg.norm(x - ref, axis=1)
    den = np.linalg.norm(ref, axis=1)
    rel = np.where(den > 0, num / np.where(den > 0, den, 1), 0.0)
    ref_pk = np.max(np.abs(ref), axis=1)
    pk = np.max(np.abs(x - ref), axis=1) / np.maximum(ref_pk, 1e-300)
    reached = ref_pk >= REACHED * max(float(ref_pk.max()), 1e-300)
    r, p = rel[reached], pk[reached]
    return {"rel_l2": rel.tolist(), "peak": pk.tolist(), "ref_peak": ref_pk.tolist(),
            "reached": int(reached.sum()), "unreached": np.flatnonzero(~reached).tolist(),
            "rel_l2_max": float(r.max()), "rel_l2_median": float(np.median(r)),
            "peak_max": float(p.max()), "peak_median": float(np.median(p)),
            "rel_l2_max_all": float(rel.max())}


def study(scene, formats=("f32", "f16", "f16diff", "bf16"), device: int = 0, log=print) -> dict:
    """Run `scene` in FP64 and in each storage format; return the error report."""
    out = {"scene": scene.name, "nodes": scene.n_nodes, "steps": scene.n_steps,
           "receivers": int(len(scene.rec_xyz)), "formats": {}}
    with Solver(device) as s:
        s.load_scene(scene)
        t0 = time.perf_counter()
        s.run(scene.n_steps, N.F64)
        ref = s.responses(np.float64)
        log(f"FP64 reference: {time.perf_counter() - t0:.1f} s")
        out["fp64_peak"] = float(np.abs(ref).max())
        src_peak = float(np.abs(scene.signals).max())
        for f in formats:
            prec = FORMATS[f]
            e = 0
            if prec in (N.F16, N.BF16, N.F16_DIFF):
                e = scale_exponent(src_peak)
                s.prepare_fields(prec, e)
            t0 = time.perf_counter()
            s.run(scene.n_steps, prec)
            x = s.responses(np.float64)
            rep = errors(x, ref)
            pf = errors(post_filter(x, POST_FS, POST_FRAC), post_filter(ref, POST_FS, POST_FRAC))
            rep["post_filtered"] = {k: pf[k] for k in ("rel_l2_median", "rel_l2_max",
                                                        "peak_median", "peak_max")}
            rep["scale_exp"] = e
            rep["seconds"] = round(time.perf_counter() - t0, 2)
            out["formats"][f] = rep
            log(f"{f}: rel L2 max {rep['rel_l2_max']:.3e} median {rep['rel_l2_median']:.3e}, "
                f"peak max {rep['peak_max']:.3e} over {rep['reached']} reached receivers; "
                f"post-filtered rel L2 median {rep['post_filtered']['rel_l2_median']:.3e} "
                f"({rep['seconds']} s)")
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--scene", default="c3", choices=["c1", "c2", "c3"])
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--scale", type=int, default=1, help="divide every room dimension (c2/c3)")
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    if args.scene == "c1":
        sc = rooms.scene_c1(n_steps=args.steps or 2000)
    elif args.scene == "c2":
        sc = rooms.scene_c2(n_steps=args.steps or 4000, scale=args.scale)
    else:
        sc = rooms.scene_c3(n_steps=args.steps or 3000, scale=args.scale)
    rep = study(sc, log=lambda m: print(m, file=sys.stderr, flush=True))
    text = json.dumps(rep)
    if args.out:
        Path(args.out).parent.mkdir(parents=True, exist_ok=True)
        Path(args.out).write_text(text)
    summary = {k: {kk: v[kk] for kk in ("rel_l2_max", "rel_l2_median", "peak_max", "peak_median",
                                        "reached", "scale_exp", "post_filtered")}
               for k, v in rep["formats"].items()}
    print(json.dumps({"scene": rep["scene"], "steps": rep["steps"], "summary": summary}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
