"""Builds the native libraries in-tree.

* ``parallelfdtd_amd/libpfdtd.so`` -- the HIP kernels and the C ABI (include/pfdtd.h),
  compiled by hipcc for gfx950 only.
* ``parallelfdtd_amd/libpfdtd_app.so`` -- the FDTD::App facade (plain C++ over the C ABI).
* ``parallelfdtd_amd/libPyFDTD*.so`` -- pybind11 module with the reference's Python surface.
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

# gfx950 only. Packed-FP32 instruction selection is disabled: the compiler would otherwise pair
# ordinary float code into v_pk_* instructions, some forms of which have given wrong results
# on this part.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

LIB_SOURCES = ["pfdtd_kernels.hip", "pfdtd_fused.hip", "pfdtd_capi.hip", "pfdtd_pool.hip",
               "pfdtd_voxel.hip", "pfdtd_probe.hip"]


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


def build_app() -> Path:
    """Compile the FDTD::App facade (HIP-free C++, plain g++) into libpfdtd_app.so."""
    out = PKG / "libpfdtd_app.so"
    tmp = out.with_suffix(".so.tmp")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", str(ROOT / "include"),
          str(CSRC / "app" / "fdtd_app.cpp"), "-L", str(PKG), "-lpfdtd",
          "-Wl,-rpath,$ORIGIN", "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_pymodule() -> Path:
    """Compile the pybind11 module libPyFDTD (the reference's Python surface)."""
    import sysconfig

    import pybind11

    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = PKG / f"libPyFDTD{suffix}"
    tmp = PKG / "libPyFDTD.so.tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
          "-I", str(ROOT / "include"), "-I", str(CSRC / "app"), "-I", pybind11.get_include(),
          "-I", sysconfig.get_paths()["include"], str(CSRC / "app" / "pyfdtd.cpp"),
          "-L", str(PKG), "-lpfdtd_app", "-lpfdtd", "-Wl,-rpath,$ORIGIN", "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_oracle() -> Path:
    """Compile the CPU oracle (oracle/Makefile)."""
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "liboracle.so"


def build_all() -> None:
    build_lib()
    build_app()
    build_pymodule()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "lib": build_lib, "app": build_app, "py": build_pymodule,
         "oracle": build_oracle}[w]()
