"""Builds the native libraries in-tree.

* ``parallelfdtd_amd/libpfdtd.so`` -- the HIP kernels and the C ABI (include/pfdtd.h),
  compiled by hipcc for gfx950 only.
* ``oracle/liboracle.so`` -- the CPU parity oracle (test infrastructure), compiled by gcc.

Run ``python -m parallelfdtd_amd.build`` or call :func:`build_all`.
"""
from __future__ import annotations

import os
import shlex
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# gfx950 only; packed-FP32 instruction selection disabled ("Build").
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

LIB_SOURCES = ["pfdtd_kernels.hip", "pfdtd_fused.hip", "pfdtd_capi.hip"]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    print("+", " ".join(shlex.quote(c) for c in cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_lib(out: Path | None = None) -> Path:
    """Compile libpfdtd.so (kernels + C ABI) with hipcc."""
    out = out or PKG / "libpfdtd.so"
    objs = []
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    for src in LIB_SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        _run([HIPCC, *HIP_FLAGS, "-I", str(ROOT / "include"), "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(str(obj))
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *objs])
    os.replace(tmp, out)
    return out


def build_oracle() -> Path:
    """Compile the CPU oracle (oracle/Makefile)."""
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "liboracle.so"


def build_all() -> None:
    build_lib()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "lib": build_lib, "oracle": build_oracle}[w]()
