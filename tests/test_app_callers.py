"""The reference's C++ callers build against the facade headers (App.h, global_includes.h,
logger.h next to fdtd_app.h) with a plain C++ compiler, and link libpfdtd_app.so:

* the reference's own src/main.cpp (read from the reference tree where it is present, fed to
  the compiler on stdin so that none of the reference's headers are found), compiled and
  linked, not run;
* a translation unit making every FDTD::App call of matlab/mex_FDTD.cpp:183-365 with its
  argument types, through the "./includes/App.h" path of matlab/mex_includes.h:5;
* the logger those callers use (loggerInit, log_msg<LEVEL>(L"...") % args), compiled and run:
  printf-style directives are replaced in order, LOG_INFO goes to solver_log.txt and LOG_ERROR
  also to standard output (src/logger.h:46-66).
No GPU: nothing here calls into the device.
"""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
APP_DIR = ROOT / "parallelfdtd_amd" / "csrc" / "app"
LIB_DIR = ROOT / "parallelfdtd_amd"
REF_MAIN = Path("/root/reference/src/main.cpp")

GXX = shutil.which("g++")
pytestmark = pytest.mark.skipif(GXX is None, reason="g++ not available")


def link_args():
    if not (LIB_DIR / "libpfdtd_app.so").exists():
        pytest.fail("libpfdtd_app.so is not built (python -m parallelfdtd_amd.build)")
    return ["-L", str(LIB_DIR), "-lpfdtd_app", "-lpfdtd", f"-Wl,-rpath,{LIB_DIR}"]


@pytest.mark.skipif(not REF_MAIN.exists(), reason="reference tree not present")
def test_reference_main_builds_unchanged(tmp_path):
    src = REF_MAIN.read_bytes()
    cmd = [GXX, "-std=c++17", "-x", "c++", "-", "-I", str(APP_DIR), "-o", str(tmp_path / "main")]
    r = subprocess.run(cmd + link_args(), input=src, cwd=tmp_path, capture_output=True)
    assert r.returncode == 0, r.stderr.decode(errors="replace")[-3000:]
    assert (tmp_path / "main").exists()


def test_mex_gateway_calls_build(tmp_path):
    (tmp_path / "includes").symlink_to(APP_DIR, target_is_directory=True)
    src = ROOT / "tests" / "app_callers" / "mex_calls.cpp"
    shutil.copy(src, tmp_path / "mex_calls.cpp")
    cmd = [GXX, "-std=c++17", "-Wall", "-Werror", "-Wno-unused-variable", "-shared", "-fPIC",
           "mex_calls.cpp", "-o", "mex_calls.so"]
    r = subprocess.run(cmd + link_args(), cwd=tmp_path, capture_output=True)
    assert r.returncode == 0, r.stderr.decode(errors="replace")[-3000:]


def test_logger_formats_and_routes(tmp_path):
    prog = tmp_path / "log.cpp"
    prog.write_text(r'''
#include "global_includes.h"
#include <string>
int main() {
  loggerInit();
  log_msg<LOG_INFO>(L"Main: begin");
  log_msg<LOG_INFO>(L"step %d of %u, %.2f s, %s, 100%%") % -5 % 10u % 1.2345 % std::string("ok");
  log_msg<LOG_ERROR>(L"bad %f %d") % 0.5f % 3;
  log_msg<LOG_DEBUG>(L"not written %d") % 1;
  log_msg<LOG_INFO>(L"R %f -> Y %f") % 0.9f % reflection2Admitance(0.9f);
  return 0;
}
''')
    (tmp_path / "solver_log.txt").write_text("stale\n")
    exe = tmp_path / "log"
    r = subprocess.run([GXX, "-std=c++17", "-I", str(APP_DIR), str(prog), "-o", str(exe)]
                       + link_args(), cwd=tmp_path, capture_output=True)
    assert r.returncode == 0, r.stderr.decode(errors="replace")[-3000:]
    out = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, check=True)
    lines = (tmp_path / "solver_log.txt").read_text().splitlines()
    assert lines == ["4 Main: begin", "4 step -5 of 10, 1.23 s, ok, 100%", "2 bad 0.500000 3",
                     "4 R 0.900000 -> Y 0.052632"]
    assert out.stdout.splitlines() == ["2 bad 0.500000 3"]
