"""The C-ABI library loads and exports every entry point include/pfdtd.h declares (no compute
calls: this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

from parallelfdtd_amd import native

HEADER = Path(__file__).resolve().parent.parent / "include" / "pfdtd.h"


def declared():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(pfdtd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "pfdtd_run" in names and "pfdtd_ctx_create" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    lib = native.lib()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(declared()) <= set(native.SIGNATURES)


def test_error_path_without_device():
    lib = native.lib()
    rc = lib.pfdtd_ctx_destroy(None)
    assert rc == native.OK
    rc = lib.pfdtd_set_kernel_timing(None, 1)
    assert rc == native.E_ARG
    assert b"null" in lib.pfdtd_last_error()


def test_job_struct_layout_matches_header(tmp_path):
    """ctypes' pfdtd_job has the field offsets and size the C compiler gives the header's."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    fields = [f[0] for f in native.PfdtdJob._fields_]
    cname = {"lambda_": "lambda"}
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {",
             '  printf("size %zu\\n", sizeof(pfdtd_job));']
    for f in fields:
        lines.append(f'  printf("{f} %zu\\n", offsetof(pfdtd_job, {cname.get(f, f)}));')
    lines += ["  return 0;", "}"]
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([gcc, "-std=c99", "-o", str(exe), str(src)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                 text=True).stdout.splitlines())
    assert int(out["size"]) == ctypes.sizeof(native.PfdtdJob)
    for f in fields:
        assert int(out[f]) == getattr(native.PfdtdJob, f).offset, f


def test_pool_job_bytes_without_device():
    """Arena footprint: two fields, three byte planes, the coefficient table, source/receiver
    tables and the response buffer, each rounded up to 256 bytes."""
    lib = native.lib()
    j = native.PfdtdJob(X=10, Y=9, Z=8, n_src=3, n_samples=50, n_rec=2, n_steps=100,
                        precision=native.F16)
    b = ctypes.c_int64()
    assert lib.pfdtd_pool_job_bytes(ctypes.byref(j), ctypes.byref(b)) == native.OK
    up = lambda v: (v + 255) // 256 * 256
    N = 10 * 9 * 8
    expect = (2 * up(2 * N) + 3 * up(N) + up(16 * 2048) + up(8 * 3) + up(4 * 3) + up(4 * 150)
              + up(8 * 2) + up(8 * 200))
    assert b.value == expect
    j.precision = native.F64
    assert lib.pfdtd_pool_job_bytes(ctypes.byref(j), ctypes.byref(b)) == native.E_ARG
    assert lib.pfdtd_pool_destroy(None) == native.OK


def test_voxel_grid_without_device():
    """pfdtd_voxel_grid is host-only: dims of the reference's box room (python/Data/box.json,
    a 1 m cube) at dx = 0.1 are floor(1/0.1) + 3 = 13 per axis, padded to (32, 4, 1)."""
    from parallelfdtd_amd.solver import voxel_grid
    from parallelfdtd_amd import geometry as G
    b = G.box_mesh((0, 0, 0), (1, 1, 1))
    assert voxel_grid(b.vertices, 0.1) == (13, 13, 13)
    assert voxel_grid(b.vertices, 0.1, (32, 4, 1)) == (32, 16, 13)
    with pytest.raises(native.PfdtdError):
        voxel_grid(b.vertices, 0.0)


def header_defines():
    return {m.group(1): int(m.group(2)) for m in
            re.finditer(r"#define\s+(PFDTD_[A-Z0-9_]+)\s+\(?(-?\d+)\)?", HEADER.read_text())}


def test_binding_constants_match_header():
    """The Python binding's numeric constants are the header's."""
    d = header_defines()
    pairs = {"PFDTD_F32": native.F32, "PFDTD_F64": native.F64, "PFDTD_F16": native.F16,
             "PFDTD_BF16": native.BF16, "PFDTD_F16_DIFF": native.F16_DIFF, "PFDTD_SWEEP_AUTO": native.SWEEP_AUTO,
             "PFDTD_SWEEP_COLUMN": native.SWEEP_COLUMN, "PFDTD_SWEEP_WIDE": native.SWEEP_WIDE,
             "PFDTD_SWEEP_PAIR": native.SWEEP_PAIR,
             "PFDTD_BATCH_INTERLEAVED": native.BATCH_INTERLEAVED,
             "PFDTD_BATCH_GROUPED": native.BATCH_GROUPED,
             "PFDTD_OK": native.OK, "PFDTD_E_ARG": native.E_ARG, "PFDTD_E_HIP": native.E_HIP,
             "PFDTD_E_NOMEM": native.E_NOMEM, "PFDTD_E_STATE": native.E_STATE,
             "PFDTD_E_DOMAIN": native.E_DOMAIN, "PFDTD_SRC_HARD": native.SRC_HARD,
             "PFDTD_SRC_SOFT": native.SRC_SOFT, "PFDTD_SRC_TRANSPARENT": native.SRC_TRANSPARENT,
             "PFDTD_STATIC_CAPTURE_MAX": native.STATIC_CAPTURE_MAX}
    for name, value in pairs.items():
        assert d[name] == value, name
